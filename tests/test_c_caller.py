"""A C program compiled against include/fpe_engine.h and linked to
libfpe_engine.so calls every symbol of the reference's C ABI
(cpp/include/fpe_engine.h:87-186) in its call shapes: the host-side calls
(create, knots, grid points, free, destroy, NULL handles) run here, the
pricing calls on the GPU against the oracle. The reference's own C++ wrapper
(cpp/include/fpe_compute.hpp) is compiled against this header when the
reference tree is present."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "c_caller"))
import caller_build  # noqa: E402

HESTON = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1)
DOC = dict(HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0, option_type=2, barrier=50.0)
REF_HPP = "/root/reference/cpp/include/fpe_compute.hpp"


def _run(mode):
    exe = caller_build.OUT
    if not os.path.exists(exe):
        caller_build.build()
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=240)
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        out[parts[0]] = parts[1:]
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def _vec(out, key):
    n = int(out[key][0])
    return np.array([float(x) for x in out[key][1:1 + n]])


def test_c_caller_host_calls(oracle):
    out = _run("host")
    assert out["invalid_handle_is_null"] == ["1"] and out["invalid_barrier_is_null"] == ["1"]
    P = oracle.Pipe(**DOC)
    s, v = P.knots()
    np.testing.assert_array_equal(_vec(out, "knots_s"), s)
    np.testing.assert_array_equal(_vec(out, "knots_v"), v)
    g = P.grid()
    assert [int(x) for x in out["grid_lens"]] == [len(g["s"]), len(g["v"])]
    ws, wv = (float(x) for x in out["grid_weight_sums"])
    np.testing.assert_allclose([ws, wv], [g["sw"].sum(), g["vw"].sum()], rtol=1e-14)
    np.testing.assert_allclose([float(x) for x in out["grid_s_range"]], [g["s"][0], g["s"][-1]], rtol=1e-15)
    for k in ("freed_vec2_nulls", "null_price", "null_pdf", "negative_nk_greeks", "host_done"):
        assert out[k] == ["1"], k


@pytest.mark.skipif(not os.path.exists(REF_HPP), reason="the reference tree is not present")
def test_reference_cpp_wrapper_compiles_against_this_header(tmp_path):
    # a link in a scratch directory: the wrapper's #include "fpe_engine.h" then
    # resolves to this repository's header (the include path), not its neighbour
    (tmp_path / "fpe_compute.hpp").symlink_to(REF_HPP)
    src = tmp_path / "use.cpp"
    src.write_text('#include "fpe_compute.hpp"\nint main() { return 0; }\n')
    root = os.path.dirname(HERE)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", str(tmp_path), "-I",
                        os.path.join(root, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(REF_HPP), reason="the reference tree is not present")
def test_reference_cpp_wrapper_runs_against_this_library(tmp_path, oracle):
    # the reference's RAII class (cpp/include/fpe_compute.hpp:104-260), compiled against this header and linked to
    # libfpe_engine.so: construction, the host-side calls, the exception on invalid parameters and the move
    (tmp_path / "fpe_compute.hpp").symlink_to(REF_HPP)
    src = tmp_path / "use.cpp"
    src.write_text(r"""
#include "fpe_compute.hpp"
#include <cstdio>
#include <stdexcept>
#include <utility>
static void put(const char* key, const std::vector<double>& v) {
    std::printf("%s %zu", key, v.size());
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}
int main() {
    fpe::FpeCompute c(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 16, 16, 50.0, 2, 50, -1.0, 180.0);
    const auto k = c.knots();
    put("knots_s", k.s);
    put("knots_v", k.v);
    const auto g = c.grid_points();
    put("grid_s", g.s);
    put("grid_v", g.v);
    put("grid_sw", g.s_weights);
    put("grid_vw", g.v_weights);
    bool threw = false;
    try {
        fpe::FpeCompute bad(-1.0, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 16, 16, 50.0, 2);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    std::printf("invalid_throws %d\n", threw ? 1 : 0);
    fpe::FpeCompute moved = std::move(c);
    std::printf("moved_valid %d %d\n", moved.valid() ? 1 : 0, c.valid() ? 1 : 0);
    return 0;
}
""")
    root = os.path.dirname(HERE)
    libdir = os.path.join(root, "fpe_option_amd")
    exe = tmp_path / "use"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(tmp_path), "-I", os.path.join(root, "include"), str(src),
                        "-o", str(exe), "-L", libdir, "-lfpe_engine", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        out[parts[0]] = parts[1:]
    P = oracle.Pipe(**DOC)
    s, v = P.knots()
    np.testing.assert_array_equal(_vec(out, "knots_s"), s)
    np.testing.assert_array_equal(_vec(out, "knots_v"), v)
    g = P.grid()
    np.testing.assert_array_equal(_vec(out, "grid_s"), g["s"])
    np.testing.assert_array_equal(_vec(out, "grid_v"), g["v"])
    np.testing.assert_array_equal(_vec(out, "grid_sw"), g["sw"])
    np.testing.assert_array_equal(_vec(out, "grid_vw"), g["vw"])
    assert out["invalid_throws"] == ["1"]
    assert out["moved_valid"] == ["1", "0"]


@pytest.mark.gpu
def test_c_caller_pricing_calls(engine, oracle):
    out = _run("gpu")
    P = oracle.Pipe(**DOC)
    q0, _ = P.initial_condition()
    np.testing.assert_allclose(_vec(out, "q0"), q0, rtol=1e-9, atol=1e-12 * np.max(np.abs(q0)))
    P.solve()
    assert [int(x) for x in out["solve_shape"]] == [P.n_states, P.N]
    qT, _ = P.state(P.n_states - 1)
    assert np.max(np.abs(_vec(out, "qT") - qT)) <= 1e-8 * np.max(np.abs(qT))
    pdf = P.pdf()
    assert [int(x) for x in out["pdf_shape"]] == list(pdf.shape)
    np.testing.assert_allclose(float(out["pdf_max"][0]), pdf.max(), rtol=1e-8)
    K = [55.0, 65.0, 80.0]
    np.testing.assert_allclose(_vec(out, "price"), P.price_at(K), rtol=1e-8, atol=1e-13)
    (p, d, g, v), rc = oracle.price_greeks([60.0, 100.0], **HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0)
    tp = 1e-8 * np.max(np.abs(p)) + 1e-13
    np.testing.assert_allclose(_vec(out, "oneshot_price"), p, rtol=0, atol=tp)
    np.testing.assert_allclose(_vec(out, "oneshot_delta"), d, rtol=0, atol=tp / (0.01 * 60.0))
    np.testing.assert_allclose(_vec(out, "oneshot_gamma"), g, rtol=0, atol=4 * tp / (0.01 * 60.0) ** 2)
    np.testing.assert_allclose(_vec(out, "oneshot_vega"), v, rtol=0, atol=tp / (0.1 * 0.1))
    assert len(_vec(out, "delta")) == 3 and np.all(np.isfinite(_vec(out, "vega")))
    assert out["gpu_done"] == ["1"]
