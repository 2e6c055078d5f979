"""HIP element kernels generated from components.

The reference generates A2D C++ for each component (reverse / forward
automatic differentiation of the element Lagrangian, amigo/component.py:
835-1167, amigo/expressions.py:386-700) and compiles it with CMake
(amigo/model.py:1409-1622).  Here the component's Lagrangian expression
(``Component.lagrangian()``) is differentiated symbolically on its DAG:

    g_i = dL/dz_i,   H_ij = dg_i/dz_j  (j <= i, mirrored)

with every shared subexpression computed once, and written out as one
straight-line device function per element.  The generated file holds that
function, the element kernel (gather the inputs, evaluate, store the
gradient and the dense Hessian, per-block Lagrangian partial sums) and an
``extern "C"`` launcher with the signature of
``amigo_elements_eval_masked`` (include/amigo_assembly.h).  It is compiled
for gfx950 with hipcc into a shared library cached by the hash of its
source, and the launcher is registered with the assembly library under a
generated element type number, so element groups of that type go through
the same assembly, batching and driver paths as the built-in types.

The generated arithmetic is compiled without contraction of products into
fused multiply-adds: the gradient-only and the Hessian evaluations then give
the same gradient bit for bit.
"""

import ctypes
import hashlib
import math
import os
import shutil
import subprocess

import numpy as np

from . import expressions as ex

ARCH = "gfx950"
HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "_gen")
GENERATED_BASE = 100
MAX_GENERATED = 256


def _lit(v):
    if math.isnan(v):
        return "__builtin_nan(\"\")"
    if math.isinf(v):
        return "__builtin_inf()" if v > 0 else "(-__builtin_inf())"
    r = repr(float(v))
    return r if ("." in r or "e" in r or "n" in r) else r + ".0"


_CALL = {
    "sin": "sin", "cos": "cos", "tan": "tan", "asin": "asin", "acos": "acos", "atan": "atan",
    "sinh": "sinh", "cosh": "cosh", "tanh": "tanh", "asinh": "asinh", "acosh": "acosh",
    "atanh": "atanh", "exp": "exp", "log": "log", "log10": "log10", "sqrt": "sqrt",
    "fabs": "fabs",
}


def _expr_code(n, name):
    """C expression of node n, its arguments by name (leaf or temporary)."""
    op = n.op
    a = [name(c) for c in n.args]
    if op == "neg":
        return f"-{a[0]}"
    if op in _CALL:
        return f"{_CALL[op]}({a[0]})"
    if op == "sign":
        return f"({a[0]} > 0.0 ? 1.0 : ({a[0]} < 0.0 ? -1.0 : 0.0))"
    if op == "add":
        return f"{a[0]} + {a[1]}"
    if op == "sub":
        return f"{a[0]} - {a[1]}"
    if op == "mul":
        return f"{a[0]} * {a[1]}"
    if op == "div":
        return f"{a[0]} / {a[1]}"
    if op == "pow":
        return f"pow({a[0]}, {a[1]})"
    if op == "atan2":
        return f"atan2({a[0]}, {a[1]})"
    if op == "min2":
        return f"({a[0]} < {a[1]} ? {a[0]} : {a[1]})"
    if op == "max2":
        return f"({a[0]} < {a[1]} ? {a[1]} : {a[0]})"
    if op == "lt":
        return f"({a[0]} < {a[1]} ? {a[2]} : {a[3]})"
    raise ValueError(op)


def _leaf_name(n):
    if n.op == "const":
        return _lit(n.val)
    if n.op == "in":
        return f"z[{n.val}]"
    if n.op == "data":
        return f"d[{n.val}]"
    if n.op == "alpha":
        return "alpha"
    return None


def _postorder(roots):
    """The nodes the roots depend on, each after its arguments: a depth-first
    post-order from the roots in order, arguments left to right, so the
    order (and the generated text) depends only on the expressions, not on
    what else was built in the process."""
    out, done = [], set()
    for r in roots:
        stack = [(r, False)]
        while stack:
            n, expanded = stack.pop()
            if n.id in done:
                continue
            if expanded:
                done.add(n.id)
                out.append(n)
                continue
            stack.append((n, True))
            for a in reversed(n.args):
                if a.id not in done:
                    stack.append((a, False))
    return out


def _body(outputs, indent="    "):
    """Straight-line code computing the outputs [(target, node)]."""
    lines = []
    name_of = {}

    def name(n):
        leaf = _leaf_name(n)
        return leaf if leaf is not None else name_of[n.id]

    for n in _postorder([n for _, n in outputs]):
        if _leaf_name(n) is not None:
            continue
        t = f"t{len(name_of)}"
        lines.append(f"{indent}const double {t} = {_expr_code(n, name)};")
        name_of[n.id] = t
    for target, n in outputs:
        lines.append(f"{indent}{target} = {name(n)};")
    return lines


class ElementDerivatives:
    """L, gradient and Hessian nodes of a component's element Lagrangian."""

    def __init__(self, comp):
        self.comp = comp
        self.nc = comp.num_inputs()
        self.ndata = comp.num_data()
        L = comp.lagrangian().node
        memo = self._memo = {}
        self.L = L
        self.g = [ex.derivative(L, i, memo) for i in range(self.nc)]
        self.H = [[None] * self.nc for _ in range(self.nc)]
        for i in range(self.nc):
            for j in range(i + 1):
                h = ex.derivative(self.g[i], j, memo)
                self.H[i][j] = self.H[j][i] = h
        self.structure = np.array(
            [[not ex._is_const(self.H[i][j], 0.0) for j in range(self.nc)] for i in range(self.nc)]
        )

    def data_derivatives(self, z, d, alpha=1.0):
        """(dL/dd, dg/dd) with numpy: E x ndata and E x nc x ndata, the
        derivatives of the Lagrangian and of its gradient with respect to
        the element's data values (post-optimality sensitivities)."""
        z = np.atleast_2d(np.asarray(z, np.float64))
        d = np.atleast_2d(np.asarray(d, np.float64))
        memo = self._memo
        Ld = [ex.derivative(self.L, -(j + 1), memo) for j in range(self.ndata)]
        gd = [ex.derivative(self.g[i], -(j + 1), memo) for i in range(self.nc) for j in range(self.ndata)]
        vals = ex.evaluate(Ld + gd, z, d, alpha=alpha)
        E = z.shape[0]
        dL = np.stack(vals[:self.ndata], axis=1) if self.ndata else np.zeros((E, 0))
        dg = (np.stack(vals[self.ndata:], axis=1).reshape(E, self.nc, self.ndata) if self.ndata
              else np.zeros((E, self.nc, 0)))
        return dL, dg

    def evaluate(self, z, d, alpha=1.0):
        """(L, g, H) with numpy for inputs z (E x nc) and data d (E x ndata)."""
        z = np.atleast_2d(np.asarray(z, np.float64))
        d = np.atleast_2d(np.asarray(d, np.float64)) if self.ndata else np.zeros((z.shape[0], 1))
        nodes = [self.L] + self.g + [self.H[i][j] for i in range(self.nc) for j in range(self.nc)]
        vals = ex.evaluate(nodes, z, d, alpha=alpha)
        E = z.shape[0]
        L = vals[0]
        g = np.stack(vals[1:1 + self.nc], axis=1)
        H = np.stack(vals[1 + self.nc:], axis=1).reshape(E, self.nc, self.nc)
        return L, g, H


def generate_source(comp):
    """(HIP source, ElementDerivatives) of a component's element kernel."""
    der = ElementDerivatives(comp)
    nc = der.nc
    lg = [("L", der.L)] + [(f"g[{i}]", der.g[i]) for i in range(nc)]
    full = list(lg)
    for i in range(nc):
        for j in range(i + 1):
            if der.structure[i, j]:
                full.append((f"H[{i * nc + j}]", der.H[i][j]))
    mirror = [f"    H[{j * nc + i}] = H[{i * nc + j}];"
              for i in range(nc) for j in range(i) if der.structure[i, j]]
    zero = [f"    H[{i * nc + j}] = 0.0;" for i in range(nc) for j in range(nc)
            if not der.structure[i, j]]
    src = f"""// Element kernel of component {comp.name}, generated by amigo_amd.codegen
// from the symbolic derivatives of its element Lagrangian.
// Inputs per element: {nc} ({comp.inputs.size} inputs, {comp.constraints.size} multipliers);
// data values per element: {der.ndata}.
#include <hip/hip_runtime.h>

#include <cmath>

namespace {{

constexpr int NT = 256;
constexpr int NC = {nc};
constexpr int PAD = -2;  // first input of a padding element

#pragma clang fp contract(off)
__device__ __forceinline__ void eval_g(const double* __restrict__ z, const double* __restrict__ d,
                                       double alpha, double& L, double* g) {{
    (void)z; (void)d; (void)alpha;
{chr(10).join(_body(lg))}
}}

__device__ __forceinline__ void eval_h(const double* __restrict__ z, const double* __restrict__ d,
                                       double alpha, double& L, double* g, double* H) {{
    (void)z; (void)d; (void)alpha;
{chr(10).join(_body(full))}
{chr(10).join(mirror)}
{chr(10).join(zero)}
}}
#pragma clang fp contract(fast)

template <bool WANT_H>
__global__ __launch_bounds__(NT) void element_kernel(int E, const int* __restrict__ idx,
                                                     const double* __restrict__ data, int ndata,
                                                     const double* __restrict__ x, double alpha,
                                                     const double* __restrict__ alpha_blk,
                                                     double* __restrict__ gbuf,
                                                     double* __restrict__ hbuf,
                                                     double* __restrict__ lpart,
                                                     const int* __restrict__ blk_grp,
                                                     const int* __restrict__ grp_active) {{
  __shared__ double red[NT / 64];
  if (blk_grp && !grp_active[blk_grp[blockIdx.x]]) return;
  const int e = blockIdx.x * NT + threadIdx.x;
  if (alpha_blk) alpha = alpha_blk[blockIdx.x];
  double Lval = 0.0;
  if (e < E && idx[(size_t)e * NC] != PAD) {{
    double z[NC], g[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {{
      const int k = idx[(size_t)e * NC + q];
      z[q] = k >= 0 ? x[k] : 0.0;
    }}
    const double* d = data + (size_t)e * ndata;
    if constexpr (WANT_H) {{
      double H[NC * NC];
      eval_h(z, d, alpha, Lval, g, H);
#pragma unroll
      for (int q = 0; q < NC * NC; q++) hbuf[(size_t)e * NC * NC + q] = H[q];
    }} else {{
      eval_g(z, d, alpha, Lval, g);
    }}
    if (gbuf) {{
#pragma unroll
      for (int q = 0; q < NC; q++) gbuf[(size_t)e * NC + q] = g[q];
    }}
  }}
  if (lpart) {{
    double v = Lval;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {{
      double s = 0.0;
      for (int w = 0; w < NT / 64; w++) s += red[w];
      lpart[blockIdx.x] = s;
    }}
  }}
}}

}}  // namespace

extern "C" int amigo_gen_num_inputs(void) {{ return NC; }}

extern "C" int amigo_gen_eval(int E, const int* idx, const double* data, int ndata,
                              const double* x, double alpha, const double* alpha_blk,
                              const int* blk_grp, const int* grp_active, double* gbuf,
                              double* hbuf, double* lpart, void* stream) {{
  if (E <= 0) return 0;
  const int nb = (E + NT - 1) / NT;
  hipStream_t s = (hipStream_t)stream;
  if (hbuf)
    hipLaunchKernelGGL(element_kernel<true>, dim3(nb), dim3(NT), 0, s, E, idx, data, ndata, x,
                       alpha, alpha_blk, gbuf, hbuf, lpart, blk_grp, grp_active);
  else
    hipLaunchKernelGGL(element_kernel<false>, dim3(nb), dim3(NT), 0, s, E, idx, data, ndata, x,
                       alpha, alpha_blk, gbuf, hbuf, lpart, blk_grp, grp_active);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}}
"""
    return src, der


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: generated element kernels need the HIP toolchain")
    return exe


def compile_source(src, out):
    """Compile a generated source into the shared library `out`."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cpp = out[:-3] + ".hip"
    with open(cpp, "w") as f:
        f.write(src)
    tmp = out + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", cpp, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {cpp}\n{r.stderr}")
    os.replace(tmp, out)
    return out


class GeneratedElement:
    """A component compiled to an element kernel and registered as element
    type ``etype`` (NUM_INPUTS and STRUCTURE of amigo_amd.problem know it)."""

    def __init__(self, comp, cache_dir=None, build=True, structure=None):
        self.comp = comp
        self.source, self.derivatives = generate_source(comp)
        self.nc = self.derivatives.nc
        self.ndata = self.derivatives.ndata
        # The KKT pattern entries of an element: the Hessian's structural
        # nonzeros, or a given superset of them (e.g. another kernel's mask,
        # so that both give the same KKT pattern and ordering).
        self.structure = self.derivatives.structure.copy()
        if structure is not None:
            structure = np.asarray(structure, bool)
            if structure.shape != self.structure.shape or np.any(self.structure & ~structure):
                raise ValueError("the given structure must contain the Hessian's nonzeros")
            self.structure = structure.copy()
        digest = hashlib.sha1(self.source.encode()).hexdigest()[:16]
        self.path = os.path.join(cache_dir or CACHE, f"{comp.name}_{digest}.so")
        if not os.path.exists(self.path):
            if not build:
                raise FileNotFoundError(f"{self.path} is not built (generate it with build=True)")
            compile_source(self.source, self.path)
        self.etype = None
        self._lib = None

    def register(self):
        """Load the kernel library and register it with the assembly library."""
        from . import _lib
        from . import problem as P

        if self.etype is not None:
            return self.etype
        key = (os.path.basename(self.path), self.structure.tobytes())
        if key in _REGISTERED:
            self.etype, self._lib, _ = _REGISTERED[key]
        else:
            self._lib = ctypes.CDLL(self.path)
            if self._lib.amigo_gen_num_inputs() != self.nc:
                raise RuntimeError(f"{self.path}: input count differs from the component")
            etype = GENERATED_BASE + len(_REGISTERED)
            if etype >= GENERATED_BASE + MAX_GENERATED:
                raise RuntimeError("too many generated element types")
            fn = ctypes.cast(self._lib.amigo_gen_eval, ctypes.c_void_p).value
            _lib.check(_lib.lib().amigo_elem_register(etype, self.nc, ctypes.c_void_p(fn)), "asm")
            _REGISTERED[key] = (etype, self._lib, self)
            self.etype = etype
        if self.etype in P.STRUCTURE and not np.array_equal(P.STRUCTURE[self.etype], self.structure):
            raise ValueError(f"element type {self.etype} is registered with another structure")
        P.NUM_INPUTS[self.etype] = self.nc
        P.STRUCTURE[self.etype] = self.structure.copy()
        return self.etype


_REGISTERED = {}


def registered():
    """The generated element types registered in this process."""
    return [ge for _, _, ge in _REGISTERED.values()]


def element_type(comp, cache_dir=None, build=True, structure=None):
    """Generate, compile (once per source) and register the element kernel of
    a component; returns the GeneratedElement (its ``etype`` is the element
    type number for ElementGroup)."""
    ge = GeneratedElement(comp, cache_dir=cache_dir, build=build, structure=structure)
    ge.register()
    return ge


def build_all(comps, cache_dir=None):
    """Generate and compile the kernels of several components (in parallel);
    every call compiles again."""
    import concurrent.futures as cf

    jobs = []
    for c in comps:
        src, _ = generate_source(c)
        digest = hashlib.sha1(src.encode()).hexdigest()[:16]
        jobs.append((src, os.path.join(cache_dir or CACHE, f"{c.name}_{digest}.so")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        return list(ex.map(lambda j: compile_source(*j), jobs))
