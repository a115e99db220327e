"""`import fpe_engine` entry point: the reference package name, served by fpe_option_amd."""
from fpe_option_amd import *  # noqa: F401,F403
from fpe_option_amd import _OPTION_TYPES, __version__  # noqa: F401
