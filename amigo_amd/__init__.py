"""amigo_amd: MI355X-native backend for Amigo's interior-point KKT path.

Host symbolic analysis in C++, numeric multifrontal LDL^T, solves, IPM
vector kernels and KKT assembly as gfx950 HIP kernels behind the C ABI in
include/amigo_amd.h.
"""

__version__ = "0.1.0"
