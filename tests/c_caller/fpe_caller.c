/* A C caller of libfpe_engine.so through include/fpe_engine.h, in the call
 * shapes of the reference's C ABI (cpp/include/fpe_engine.h:87-186): every
 * reference symbol is called, results are read and released with the
 * matching free call.
 *
 *   fpe_caller host   create / knots / grid points / free / destroy, NULL
 *                     handles and invalid parameters (no GPU work)
 *   fpe_caller gpu    initial condition, solve, PDF, prices, Greeks, one-shot
 *
 * Output: one "key value ..." line per result, read by tests/test_c_caller.py.
 */
#include <stdio.h>
#include <string.h>

#include "fpe_engine.h"

static void print_vec(const char* key, const double* x, int n) {
    printf("%s %d", key, n);
    for (int i = 0; i < n; ++i) printf(" %.17g", x[i]);
    printf("\n");
}

static FpeCompute* make_doc(void) {
    /* down-and-out call, B = 50: kappa theta sigma rho r T S0 V0, n_s n_v, barrier, type, num_insert, s_min, s_max */
    return fpe_compute_create(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 16, 16, 50.0, 2, 50, 0.0, 180.0);
}

static int host_part(void) {
    FpeCompute* bad = fpe_compute_create(-1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 16, 16, 0.0, 8, 50, 0.0, 180.0);
    printf("invalid_handle_is_null %d\n", bad == NULL);
    FpeCompute* badb = fpe_compute_create(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 16, 16, 70.0, 2, 50, 0.0, 180.0);
    printf("invalid_barrier_is_null %d\n", badb == NULL);

    FpeCompute* h = make_doc();
    if (!h) {
        printf("create_failed %s\n", fpe_last_error());
        return 1;
    }
    struct FpeVec2Result k;
    fpe_compute_knots(h, &k);
    print_vec("knots_s", k.s_data, k.s_len);
    print_vec("knots_v", k.v_data, k.v_len);
    fpe_compute_free_vec2(&k);
    printf("freed_vec2_nulls %d\n", k.s_data == NULL && k.v_data == NULL);

    struct FpeGridPtsResult g;
    fpe_compute_grid_points(h, &g);
    printf("grid_lens %d %d\n", g.s_len, g.v_len);
    double ws = 0.0, wv = 0.0;
    for (int i = 0; i < g.s_len; ++i) ws += g.sw_data[i];
    for (int i = 0; i < g.v_len; ++i) wv += g.vw_data[i];
    printf("grid_weight_sums %.17g %.17g\n", ws, wv);
    printf("grid_s_range %.17g %.17g\n", g.s_data[0], g.s_data[g.s_len - 1]);
    fpe_compute_free_grid_pts(&g);

    /* NULL handle: {NULL, 0} results (c_abi.mojo:178-185) */
    struct FpeVecResult v;
    double K[2] = {55.0, 65.0};
    fpe_compute_price(NULL, K, 2, &v);
    printf("null_price %d\n", v.data == NULL && v.len == 0);
    fpe_compute_free_vec(&v);
    struct FpeMatResult m;
    fpe_compute_pdf(NULL, &m);
    printf("null_pdf %d\n", m.data == NULL && m.n_rows == 0);
    fpe_compute_free_mat(&m);
    struct FpeGreeksResult gr;
    fpe_compute_greeks(h, K, -1, 0.01, 0.1, &gr); /* n_K < 0 */
    printf("negative_nk_greeks %d\n", gr.delta == NULL && gr.len == 0);
    fpe_compute_free_greeks(&gr);
    struct FpeOneshotResult one;
    memset(&one, 0, sizeof one);
    fpe_compute_free_oneshot(&one); /* freeing an empty result is a no-op */
    fpe_compute_destroy(h);
    fpe_compute_destroy(NULL);
    printf("host_done 1\n");
    return 0;
}

static int gpu_part(void) {
    FpeCompute* h = make_doc();
    if (!h) return 1;
    struct FpeVecResult q0;
    fpe_compute_initial_condition(h, &q0);
    print_vec("q0", q0.data, q0.len);
    fpe_compute_free_vec(&q0);

    struct FpeMatResult sol;
    fpe_compute_solve(h, &sol);
    printf("solve_shape %d %d\n", sol.n_rows, sol.n_cols);
    if (sol.n_rows > 0) print_vec("qT", sol.data + (size_t)(sol.n_rows - 1) * sol.n_cols, sol.n_cols);
    fpe_compute_free_mat(&sol);

    struct FpeMatResult pdf;
    fpe_compute_pdf(h, &pdf);
    printf("pdf_shape %d %d\n", pdf.n_rows, pdf.n_cols);
    double pmax = 0.0;
    for (int i = 0; i < pdf.n_rows * pdf.n_cols; ++i) pmax = pdf.data[i] > pmax ? pdf.data[i] : pmax;
    printf("pdf_max %.17g\n", pmax);
    fpe_compute_free_mat(&pdf);

    double K[3] = {55.0, 65.0, 80.0};
    struct FpeVecResult pr;
    fpe_compute_price(h, K, 3, &pr);
    print_vec("price", pr.data, pr.len);
    fpe_compute_free_vec(&pr);

    struct FpeGreeksResult gr;
    fpe_compute_greeks(h, K, 3, 0.01, 0.1, &gr);
    print_vec("delta", gr.delta, gr.len);
    print_vec("gamma", gr.gamma, gr.len);
    print_vec("vega", gr.vega, gr.len);
    fpe_compute_free_greeks(&gr);
    fpe_compute_destroy(h);

    double K1[2] = {60.0, 100.0};
    struct FpeOneshotResult one;
    fpe_price_oneshot(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, K1, 2, 0.0, 8, 16, 16, 50, 0.0, 180.0, &one);
    print_vec("oneshot_price", one.price, one.len);
    print_vec("oneshot_delta", one.delta, one.len);
    print_vec("oneshot_gamma", one.gamma, one.len);
    print_vec("oneshot_vega", one.vega, one.len);
    fpe_compute_free_oneshot(&one);
    printf("gpu_done 1\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && strcmp(argv[1], "gpu") == 0) return gpu_part();
    return host_part();
}
