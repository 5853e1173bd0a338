"""Python front end of libmsfem_hip.so (the C ABI of include/msfem.h).

The host-side mirror of the reference's per-cell interface for Python callers:
batch_build() returns, for every coarse cell, what
DiffusionProblemBasis<dim>::get_global_element_matrix() / get_global_element_rhs()
return after run() (include/base/diffusion_problem_basis.tpp:320-334, :440-474),
plus the bases themselves.  All numerics run in the library's HIP kernels; this
module only marshals arrays.  It raises if the library is missing: there is no
CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsfem_hip.so")

OK, ERR_ARG, ERR_HIP, ERR_NOCONV, ERR_NODEVICE = 0, 1, 2, 3, 4
COEF_SCALAR_CELL, COEF_TENSOR_CELL, COEF_SCALAR_QP, COEF_TENSOR_QP = 0, 1, 2, 3
RHS_CONST, RHS_SINSIN = 0, 1
PREC_FP64, PREC_FP32_REFINE = 0, 1
BC_ZERO_DIRICHLET, BC_REFERENCE = 0, 1
N_MARKS = 8  # MSFEM_N_MARKS: mark slots per context

# Every symbol include/msfem.h declares.
EXPORTS = [
    "msfem_ctx_create", "msfem_ctx_destroy", "msfem_last_error", "msfem_default_opts", "msfem_version",
    "msfem_uses_coarse_space",
    "msfem_ctx_set_profiling", "msfem_ctx_profile", "msfem_ctx_mark", "msfem_ctx_elapsed", "msfem_ctx_aux_stream", "msfem_ctx_mark_on", "msfem_ctx_wait_mark", "msfem_dev_alloc", "msfem_dev_free", "msfem_memcpy_h2d",
    "msfem_memcpy_d2h", "msfem_ctx_stream", "msfem_ctx_synchronize", "msfem_variant_count", "msfem_set_variant",
    "msfem_batch_build", "msfem_batch_build_device", "msfem_energy_error_device", "msfem_coarse_assemble", "msfem_coarse_assemble_device",
    "msfem_coarse_solve", "msfem_coarse_solve_device", "msfem_coarse_solve_device_noalloc", "msfem_gather_weights", "msfem_reconstruct_device",
    "msfem_philox4x32_10", "msfem_gen_unit_boxes", "msfem_gen_coef_reference", "msfem_gen_coef_periodic",
    "msfem_gen_coef_lognormal", "msfem_gen_resolution_classes", "msfem_gen_coef_tensor", "msfem_gen_grf_log",
    "msfem_gen_perturbed_quads", "msfem_gen_coef_reference_quad", "msfem_gen_perturbed_hexes",
    "msfem_gen_coef_reference_hex",
    "msfem_fine_solve", "msfem_nn_gemm_bf16", "msfem_nn_gemm_bf16_splitk", "msfem_nn_im2col", "msfem_nn_conv_bf16", "msfem_nn_conv3x3_pack", "msfem_nn_conv3x3_bf16", "msfem_nn_conv3x3_bf16_ex", "msfem_nn_conv3x3_first_bf16", "msfem_nn_conv_small", "msfem_nn_head_bases", "msfem_nn_col2im", "msfem_nn_bias_act", "msfem_nn_act_bwd",
    "msfem_nn_bias_grad", "msfem_nn_pool2", "msfem_nn_up2", "msfem_nn_axpy", "msfem_nn_axpby", "msfem_nn_to_bf16",
    "msfem_nn_mse", "msfem_nn_adam", "msfem_nn_delta_to_bases", "msfem_nn_bases_to_delta",
]


class MsfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"msfem error {code}: {msg}")
        self.code = code


class SolverOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_iter", ctypes.c_int32), ("max_refine", ctypes.c_int32),
                ("partition_of_unity", ctypes.c_int32), ("rel_tol", ctypes.c_double), ("abs_tol", ctypes.c_double),
                ("inner_rel_tol", ctypes.c_double), ("variant", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("coarse_space", ctypes.c_int32), ("q1_start", ctypes.c_int32), ("energy_tol", ctypes.c_double),
                ("reserved", ctypes.c_int32 * 2)]


class Batch(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n_cells", ctypes.c_int32), ("n_fine", ctypes.c_void_p),
                ("box", ctypes.c_void_p), ("coef_kind", ctypes.c_int32), ("coef", ctypes.c_void_p),
                ("rhs_kind", ctypes.c_int32), ("rhs_value", ctypes.c_double), ("x0", ctypes.c_void_p),
                ("vertices", ctypes.c_void_p)]


class BatchOut(ctypes.Structure):
    _fields_ = [("Ae", ctypes.c_void_p), ("Fe", ctypes.c_void_p), ("bases", ctypes.c_void_p),
                ("iters", ctypes.c_void_p), ("status", ctypes.c_void_p), ("resid", ctypes.c_void_p),
                ("energy_bound", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run build.py)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.msfem_last_error.restype = ctypes.c_char_p
        L.msfem_last_error.argtypes = [vp]
        L.msfem_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
        L.msfem_ctx_destroy.argtypes = [vp]
        L.msfem_ctx_set_profiling.argtypes = [vp, i32]
        L.msfem_ctx_profile.argtypes = [vp, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.msfem_ctx_mark.argtypes = [vp, ctypes.c_int]
        L.msfem_ctx_elapsed.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(dbl)]
        L.msfem_ctx_aux_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.msfem_ctx_mark_on.argtypes = [vp, ctypes.c_int, vp]
        L.msfem_ctx_wait_mark.argtypes = [vp, ctypes.c_int, vp]
        L.msfem_dev_alloc.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.msfem_dev_free.argtypes = [vp, vp]
        L.msfem_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_uint64]
        L.msfem_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_uint64]
        L.msfem_ctx_stream.argtypes = [vp]
        L.msfem_ctx_stream.restype = vp
        L.msfem_ctx_synchronize.argtypes = [vp]
        L.msfem_variant_count.argtypes = [i32, i32]
        L.msfem_set_variant.argtypes = [vp, i32, i32, i32, i32, i32]
        L.msfem_default_opts.argtypes = [ctypes.POINTER(SolverOpts)]
        L.msfem_default_opts.restype = None
        L.msfem_uses_coarse_space.argtypes = [vp, i32, i32, ctypes.POINTER(SolverOpts), i32, i32]
        L.msfem_batch_build.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(SolverOpts), ctypes.POINTER(BatchOut)]
        L.msfem_batch_build_device.argtypes = [vp, ctypes.POINTER(Batch), ctypes.POINTER(SolverOpts),
                                               ctypes.POINTER(BatchOut), vp]
        L.msfem_energy_error_device.argtypes = [vp, ctypes.POINTER(Batch), vp, vp, vp, vp]
        L.msfem_coarse_assemble.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, vp, vp]
        L.msfem_coarse_assemble_device.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, vp, vp, vp]
        L.msfem_coarse_solve.argtypes = [vp, i32, i32, vp, vp, vp, vp, dbl, i32, vp, vp]
        L.msfem_coarse_solve_device.argtypes = [vp, i32, i32, vp, vp, vp, vp, dbl, i32, vp, vp, vp]
        L.msfem_coarse_solve_device_noalloc.argtypes = [vp, i32, i32, vp, vp, vp, vp, dbl, i32, vp, vp, vp]
        L.msfem_gather_weights.argtypes = [i32, i32, vp, vp]
        L.msfem_reconstruct_device.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp]
        L.msfem_philox4x32_10.argtypes = [vp, vp, vp]
        L.msfem_philox4x32_10.restype = None
        L.msfem_gen_unit_boxes.argtypes = [i32, i32, vp]
        L.msfem_gen_coef_reference.argtypes = [i32, i32, vp, vp, vp]
        L.msfem_gen_coef_periodic.argtypes = [i32, vp, vp, dbl, vp]
        L.msfem_gen_coef_lognormal.argtypes = [i32, i32, vp, ctypes.c_uint64, dbl, vp]
        L.msfem_gen_resolution_classes.argtypes = [i32, ctypes.c_uint64, dbl, vp, vp, vp]
        L.msfem_gen_coef_tensor.argtypes = [i32, vp, ctypes.c_uint64, dbl, vp]
        L.msfem_gen_grf_log.argtypes = [i32, i32, ctypes.c_uint64, dbl, dbl, vp]
        L.msfem_gen_perturbed_quads.argtypes = [i32, dbl, ctypes.c_uint64, vp]
        L.msfem_gen_coef_reference_quad.argtypes = [i32, vp, vp, vp]
        L.msfem_gen_perturbed_hexes.argtypes = [i32, dbl, ctypes.c_uint64, vp]
        L.msfem_gen_coef_reference_hex.argtypes = [i32, vp, vp, vp]
        L.msfem_fine_solve.argtypes = [vp, i32, i32, i32, vp, i32, dbl, i32, dbl, i32, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().msfem_last_error(ctx)
        raise MsfemError(rc, msg.decode() if msg else "")


def default_opts(**kw):
    o = SolverOpts()
    lib().msfem_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def coef_stride(dim, kind):
    nq, ns = 1 << dim, (3 if dim == 2 else 6)
    return {COEF_SCALAR_CELL: 1, COEF_TENSOR_CELL: ns, COEF_SCALAR_QP: nq, COEF_TENSOR_QP: nq * ns}[kind]


def bases_size(dim, n_fine):
    return int(sum((1 << dim) * (int(k) + 1) ** dim for k in n_fine))


class Context:
    """One GPU context (DiffusionProblemMultiscale's per-rank state)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().msfem_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().msfem_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_variant(self, dim, nf, precision, coef_kind, variant):
        """Launch-shape id for one class (0 = built-in default)."""
        _check(lib().msfem_set_variant(self.h, dim, nf, precision, coef_kind, variant), self.h)

    def load_variant_table(self, path=None):
        """Applies a table written by autotune.py (default: autotune_table.json here)."""
        import json
        path = path or os.path.join(HERE, "autotune_table.json")
        if not os.path.exists(path):
            return 0
        table = json.load(open(path))
        for e in table["entries"]:
            self.set_variant(e["dim"], e["nf"], e["precision"], e["coef_kind"], e["variant"])
        return len(table["entries"])

    def uses_coarse_space(self, dim, nf, opts, coef_kind, general=False):
        """0: a class of nf fine cells per side runs Jacobi-PCG here; 1: the
        two-level local solve; 2: the two-level solve with the intermediate hat
        levels."""
        return int(lib().msfem_uses_coarse_space(self.h, dim, nf, ctypes.byref(opts), coef_kind, 1 if general else 0))

    def set_profiling(self, on=True):
        _check(lib().msfem_ctx_set_profiling(self.h, 1 if on else 0), self.h)

    def profile(self):
        """(ms of the local solves, ms of the element matrices) of the last build."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().msfem_ctx_profile(self.h, ctypes.byref(a), ctypes.byref(b)), self.h)
        return a.value, b.value

    def mark(self, slot):
        """Record timing mark `slot` (a HIP event) on the context stream."""
        _check(lib().msfem_ctx_mark(self.h, slot), self.h)

    def mark_on(self, slot, stream):
        """Record mark `slot` on `stream` (None: the context stream)."""
        _check(lib().msfem_ctx_mark_on(self.h, slot, stream), self.h)

    def wait_mark(self, slot, stream):
        """Make `stream` (None: the context stream) wait for the last recording of mark `slot`."""
        _check(lib().msfem_ctx_wait_mark(self.h, slot, stream), self.h)

    @property
    def aux_stream(self):
        """The context's second stream (created on first use)."""
        s = ctypes.c_void_p()
        _check(lib().msfem_ctx_aux_stream(self.h, ctypes.byref(s)), self.h)
        return s.value

    def elapsed(self, start, end):
        """Device milliseconds between two recorded marks (waits for `end`)."""
        ms = ctypes.c_double()
        _check(lib().msfem_ctx_elapsed(self.h, start, end, ctypes.byref(ms)), self.h)
        return ms.value

    # ---- device memory --------------------------------------------------------------
    def alloc(self, shape, dtype):
        """A DeviceArray of the given shape/dtype in HBM (freed with the array)."""
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def stream(self):
        return lib().msfem_ctx_stream(self.h)

    def synchronize(self):
        _check(lib().msfem_ctx_synchronize(self.h), self.h)

    # ---- the hot path, host arrays ------------------------------------------------
    def batch_build(self, dim, n_fine, box, coef_kind, coef, rhs_kind=RHS_CONST, rhs_value=1.0, x0=None,
                    opts=None, want_bases=True, raise_noconv=True, vertices=None, energy_bound=False):
        """vertices: general cells, [n][4][2] quadrilaterals (2D) or [n][8][3] hexahedra (3D);
        box is then ignored and may be None.
        energy_bound: also return the a posteriori bound of each basis's relative
        energy-norm error ("energy_bound", [n][nb])."""
        n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
        n = int(n_fine.shape[0])
        nb = 1 << dim
        box = None if box is None else np.ascontiguousarray(box, dtype=np.float64)
        vertices = None if vertices is None else np.ascontiguousarray(vertices, dtype=np.float64)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        opts = opts or default_opts()
        out = {"Ae": np.zeros((n, nb, nb)), "Fe": np.zeros((n, nb)), "iters": np.zeros((n, nb), np.int32),
               "status": np.zeros((n, nb), np.int32), "resid": np.zeros((n, nb)),
               "bases": np.zeros(bases_size(dim, n_fine)) if want_bases else None,
               "energy_bound": np.zeros((n, nb)) if energy_bound else None}
        b = Batch(dim, n, _p(n_fine), _p(box), coef_kind, _p(coef), rhs_kind, rhs_value, _p(x0), _p(vertices))
        o = BatchOut(_p(out["Ae"]), _p(out["Fe"]), _p(out["bases"]), _p(out["iters"]), _p(out["status"]),
                     _p(out["resid"]), _p(out["energy_bound"]))
        rc = lib().msfem_batch_build(self.h, ctypes.byref(b), ctypes.byref(opts), ctypes.byref(o))
        if rc == ERR_NOCONV and not raise_noconv:
            return out
        _check(rc, self.h)
        return out

    # ---- the hot path, device pointers (e.g. torch tensors' data_ptr()) --------------
    def batch_build_device(self, dim, n_fine, box_ptr, coef_kind, coef_ptr, Ae_ptr, Fe_ptr, bases_ptr=None,
                           iters_ptr=None, status_ptr=None, resid_ptr=None, rhs_kind=RHS_CONST, rhs_value=1.0,
                           x0_ptr=None, opts=None, stream=None, vertices_ptr=None, energy_bound_ptr=None):
        n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
        opts = opts or default_opts()
        b = Batch(dim, int(n_fine.shape[0]), _p(n_fine), box_ptr, coef_kind, coef_ptr, rhs_kind, rhs_value, x0_ptr,
                  vertices_ptr)
        o = BatchOut(Ae_ptr, Fe_ptr, bases_ptr, iters_ptr, status_ptr, resid_ptr, energy_bound_ptr)
        _check(lib().msfem_batch_build_device(self.h, ctypes.byref(b), ctypes.byref(opts), ctypes.byref(o), stream),
               self.h)

    def energy_error_device(self, dim, n_fine, box_ptr, coef_kind, coef_ptr, bases_ptr, ref_ptr, err_ptr,
                            stream=None, vertices_ptr=None):
        """err[c][i] = ||phi_i - phi_ref,i||_K / ||phi_ref,i||_K for every basis (device pointers)."""
        n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
        b = Batch(dim, int(n_fine.shape[0]), _p(n_fine), box_ptr, coef_kind, coef_ptr, RHS_CONST, 0.0, None,
                  vertices_ptr)
        _check(lib().msfem_energy_error_device(self.h, ctypes.byref(b), bases_ptr, ref_ptr, err_ptr, stream), self.h)

    # ---- coarse system -----------------------------------------------------------------
    def coarse_assemble(self, dim, nc, Ae, Fe, bc_kind):
        nn = (nc + 1) ** dim
        Kc = np.zeros((nn, 3 ** dim))
        Fc = np.zeros(nn)
        cm = np.zeros(nn, np.uint8)
        cv = np.zeros(nn)
        Ae = np.ascontiguousarray(Ae, dtype=np.float64)
        Fe = np.ascontiguousarray(Fe, dtype=np.float64)
        _check(lib().msfem_coarse_assemble(self.h, dim, nc, _p(Ae), _p(Fe), bc_kind, _p(Kc), _p(Fc), _p(cm), _p(cv)),
               self.h)
        return Kc, Fc, cm, cv

    def coarse_solve(self, dim, nc, Kc, Fc, cmask, cval, rel_tol=1e-14, max_iter=100000):
        u = np.zeros((nc + 1) ** dim)
        it = np.zeros(1, np.int32)
        Kc = np.ascontiguousarray(Kc, dtype=np.float64)
        Fc = np.ascontiguousarray(Fc, dtype=np.float64)
        cmask = np.ascontiguousarray(cmask, dtype=np.uint8)
        cval = np.ascontiguousarray(cval, dtype=np.float64)
        _check(lib().msfem_coarse_solve(self.h, dim, nc, _p(Kc), _p(Fc), _p(cmask), _p(cval), rel_tol, max_iter,
                                        _p(u), _p(it)), self.h)
        return u, int(it[0])

    def coarse_assemble_device(self, dim, nc, Ae_ptr, Fe_ptr, bc_kind, Kc_ptr, Fc_ptr, cm_ptr, cv_ptr, stream=None):
        _check(lib().msfem_coarse_assemble_device(self.h, dim, nc, Ae_ptr, Fe_ptr, bc_kind, Kc_ptr, Fc_ptr, cm_ptr,
                                                  cv_ptr, stream), self.h)

    def coarse_solve_device(self, dim, nc, Kc_ptr, Fc_ptr, cm_ptr, cv_ptr, rel_tol, max_iter, u_ptr, it_ptr,
                            stream=None):
        _check(lib().msfem_coarse_solve_device(self.h, dim, nc, Kc_ptr, Fc_ptr, cm_ptr, cv_ptr, rel_tol, max_iter,
                                               u_ptr, it_ptr, stream), self.h)

    def reconstruct_device(self, dim, n_fine, bases_ptr, w_ptr, u_ptr, stream=None):
        n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
        _check(lib().msfem_reconstruct_device(self.h, dim, int(n_fine.shape[0]), _p(n_fine), bases_ptr, w_ptr, u_ptr,
                                              stream), self.h)

    # ---- fine-scale standard FEM (the reference's DiffusionProblem<2>) -----------------
    def fine_solve(self, n, coef_kind, coef, rhs_kind=RHS_CONST, rhs_value=1.0, bc_kind=BC_ZERO_DIRICHLET,
                   rel_tol=1e-12, max_iter=200000, dim=2):
        """Q1 solve on the n x n fine mesh of the unit square -> ((n+1)^2 nodal values, iterations)."""
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        u = np.zeros((n + 1) ** dim)
        it = ctypes.c_int32()
        _check(lib().msfem_fine_solve(self.h, dim, n, coef_kind, _p(coef), rhs_kind, rhs_value, bc_kind, rel_tol,
                                      max_iter, _p(u), ctypes.byref(it)), self.h)
        return u, it.value


class DeviceArray:
    """A buffer in device memory with a numpy shape/dtype (host copies on demand)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().msfem_dev_alloc(ctx.h, self.nbytes, ctypes.byref(p)), ctx.h)
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().msfem_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes), self.ctx.h)

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().msfem_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), self.ctx.h)
        return out

    def read(self, start, count):
        """Host copy of elements start .. start+count of the flattened buffer."""
        total = self.nbytes // self.dtype.itemsize
        if start < 0 or count < 0 or start + count > total:
            raise ValueError(f"range {start}+{count} outside a buffer of {total} elements")
        out = np.empty(count, self.dtype)
        if count:
            _check(lib().msfem_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr + start * self.dtype.itemsize,
                                          count * self.dtype.itemsize), self.ctx.h)
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().msfem_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def variant_count(dim, nf):
    return lib().msfem_variant_count(dim, nf)


def gather_weights(dim, nc, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.zeros((nc ** dim, 1 << dim))
    _check(lib().msfem_gather_weights(dim, nc, _p(u), _p(w)))
    return w


# ---- synthetic inputs (host) -------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().msfem_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def unit_boxes(dim, nc):
    box = np.zeros((nc ** dim, dim, 2))
    _check(lib().msfem_gen_unit_boxes(dim, nc, _p(box)))
    return box


def _tile_total(dim, n_fine, kind):
    return int(sum(int(k) ** dim for k in n_fine)) * coef_stride(dim, kind)


def coef_reference(dim, n_fine, box):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    box = np.ascontiguousarray(box, dtype=np.float64)
    out = np.zeros(_tile_total(dim, n_fine, COEF_TENSOR_QP))
    _check(lib().msfem_gen_coef_reference(dim, len(n_fine), _p(n_fine), _p(box), _p(out)))
    return out


def perturbed_quads(nc, amp, seed):
    """[nc*nc][4][2] vertices of the unit square's nc x nc mesh with moved interior vertices."""
    out = np.zeros((nc * nc, 4, 2))
    _check(lib().msfem_gen_perturbed_quads(nc, amp, seed, _p(out)))
    return out


def coef_reference_quad(n_fine, vertices):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    out = np.zeros(_tile_total(2, n_fine, COEF_TENSOR_QP))
    _check(lib().msfem_gen_coef_reference_quad(len(n_fine), _p(n_fine), _p(vertices), _p(out)))
    return out


def perturbed_hexes(nc, amp, seed):
    """[nc^3][8][3] vertices of the unit cube's nc^3 mesh with moved interior vertices (amp < 1/6)."""
    out = np.zeros((nc ** 3, 8, 3))
    _check(lib().msfem_gen_perturbed_hexes(nc, amp, seed, _p(out)))
    return out


def coef_reference_hex(n_fine, vertices):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    out = np.zeros(_tile_total(3, n_fine, COEF_TENSOR_QP))
    _check(lib().msfem_gen_coef_reference_hex(len(n_fine), _p(n_fine), _p(vertices), _p(out)))
    return out


def coef_periodic(n_fine, box, eps):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    box = np.ascontiguousarray(box, dtype=np.float64)
    out = np.zeros(_tile_total(2, n_fine, COEF_SCALAR_QP))
    _check(lib().msfem_gen_coef_periodic(len(n_fine), _p(n_fine), _p(box), eps, _p(out)))
    return out


def coef_lognormal(dim, n_fine, seed, clip):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    out = np.zeros(_tile_total(dim, n_fine, COEF_SCALAR_CELL))
    _check(lib().msfem_gen_coef_lognormal(dim, len(n_fine), _p(n_fine), seed, clip, _p(out)))
    return out


def resolution_classes(nc, seed, sigma, classes=(16, 32, 64, 128)):
    cl = np.ascontiguousarray(classes, dtype=np.int32)
    nf = np.zeros(nc * nc, np.int32)
    con = np.zeros(nc * nc)
    _check(lib().msfem_gen_resolution_classes(nc, seed, sigma, _p(cl), _p(nf), _p(con)))
    return nf, con


def coef_tensor(n_fine, seed, sigma):
    n_fine = np.ascontiguousarray(n_fine, dtype=np.int32)
    out = np.zeros(_tile_total(2, n_fine, COEF_TENSOR_CELL))
    _check(lib().msfem_gen_coef_tensor(len(n_fine), _p(n_fine), seed, sigma, _p(out)))
    return out


def grf_log(n_cells, nf, seed, corr_len, clip):
    out = np.zeros((n_cells, nf * nf), np.float32)
    _check(lib().msfem_gen_grf_log(n_cells, nf, seed, corr_len, clip, _p(out)))
    return out
