"""ctypes front end of the CPU FP64 oracle (msfem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the oracle.  The product library never
imports this module.

Every function restates one piece of the reference (file:line citations are in
msfem_oracle.c):
  basis_q1_coeffs / basis_q1_value   include/coefficients/basis_q1.tpp:26-113
  matrix_coeff_ref                   include/coefficients/matrix_coeff.tpp:17-91
  fine_element / local_assemble      include/base/diffusion_problem_basis.tpp:161-242
  local_bases_direct                 diffusion_problem_basis.tpp:450-465 (exact solve)
  local_bases_ssor_pcg               diffusion_problem_basis.tpp:295-317 (the recipe)
  element                            diffusion_problem_basis.tpp:247-285
  cell_build / batch_build           diffusion_problem_basis.tpp:440-474,
                                     diffusion_problem_ms.tpp:42-88
  coarse_assemble / coarse_solve     diffusion_problem_ms.tpp:108-296
  reconstruct                        diffusion_problem_basis.tpp:354-377
  *_quad (general quadrilaterals)    diffusion_problem_basis.tpp:92-99, :161-242 with the
                                     bilinear mapping; basis_q1.tpp:26-46 at mapped nodes
  *_hex (general hexahedra)          the same in 3D with the trilinear mapping
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsfem_oracle.so")

COEF_SCALAR_CELL, COEF_TENSOR_CELL, COEF_SCALAR_QP, COEF_TENSOR_QP = 0, 1, 2, 3
RHS_CONST, RHS_SINSIN = 0, 1
SOLVER_DIRECT, SOLVER_SSOR_PCG = 0, 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ubyte)
_i, _d = ctypes.c_int, ctypes.c_double


def _declare(L):
    sig = {
        "orc_n_sym": (_i, [_i]),
        "orc_coef_stride": (_i, [_i, _i]),
        "orc_basis_q1_coeffs": (_i, [_i, _dp, _dp]),
        "orc_basis_q1_value": (_d, [_i, _dp, _i, _dp]),
        "orc_box_vertices": (None, [_i, _dp, _dp]),
        "orc_matrix_coeff_ref": (None, [_i, _dp, _dp]),
        "orc_fine_element": (None, [_i, _dp, _dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_n_nodes": (_i, [_i, _i]),
        "orc_local_assemble": (None, [_i, _i, _dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_stencil_apply": (None, [_i, _i, _dp, _dp, _dp]),
        "orc_dirichlet_data": (None, [_i, _i, _dp, _i, _dp]),
        "orc_local_bases_direct": (_i, [_i, _i, _dp, _dp, _dp]),
        "orc_local_bases_ssor_pcg": (_i, [_i, _i, _dp, _dp, _d, _d, _i, _dp, _ip]),
        "orc_element": (None, [_i, _i, _dp, _dp, _dp, _dp, _dp]),
        "orc_cell_build": (_i, [_i, _i, _dp, _i, _dp, _i, _d, _i, _d, _d, _i, _dp, _dp, _dp, _ip]),
        "orc_batch_build": (_i, [_i, _i, _dp, _ip, _i, _dp, _i, _d, _i, _d, _d, _i, _i, _dp, _dp, _dp, _ip]),
        "orc_max_threads": (_i, []),
        "orc_coarse_assemble": (None, [_i, _i, _dp, _dp, _i, _dp, _dp, _up, _dp]),
        "orc_coarse_solve": (_i, [_i, _i, _dp, _dp, _up, _dp, _dp]),
        "orc_reconstruct": (None, [_i, _i, _dp, _dp, _dp]),
        "orc_quad_node": (None, [_dp, _i, _i, _i, _dp]),
        "orc_fine_element_quad": (None, [_dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_local_assemble_quad": (None, [_i, _dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_dirichlet_data_quad": (None, [_i, _dp, _i, _dp]),
        "orc_local_bases_direct_quad": (_i, [_i, _dp, _dp, _dp]),
        "orc_cell_build_quad": (_i, [_i, _dp, _i, _dp, _i, _d, _i, _d, _d, _i, _dp, _dp, _dp, _ip]),
        "orc_batch_build_quad": (_i, [_i, _dp, _ip, _i, _dp, _i, _d, _i, _d, _d, _i, _i, _dp, _dp, _dp, _ip]),
        "orc_cell_node": (None, [_i, _dp, _i, _ip, _dp]),
        "orc_fine_element_hex": (None, [_dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_local_assemble_hex": (None, [_i, _dp, _i, _dp, _i, _d, _dp, _dp]),
        "orc_dirichlet_data_hex": (None, [_i, _dp, _i, _dp]),
        "orc_local_bases_direct_hex": (_i, [_i, _dp, _dp, _dp]),
        "orc_batch_build_general": (_i, [_i, _i, _dp, _ip, _i, _dp, _i, _d, _i, _d, _d, _i, _i, _dp, _dp, _dp,
                                         _ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _d64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def n_nodes(dim, nf):
    return (nf + 1) ** dim


def coef_stride(dim, kind):
    return lib().orc_coef_stride(dim, kind)


def basis_q1_coeffs(dim, verts):
    v = _d64(verts).reshape(-1)
    c = np.zeros((1 << dim) ** 2)
    rc = lib().orc_basis_q1_coeffs(dim, _ptr(v), _ptr(c))
    if rc:
        raise ValueError("singular vertex monomial matrix")
    return c.reshape(1 << dim, 1 << dim)


def basis_q1_value(dim, coeffs, index, p):
    c = _d64(coeffs).reshape(-1)
    pp = _d64(p)
    return lib().orc_basis_q1_value(dim, _ptr(c), index, _ptr(pp))


def box_vertices(dim, box):
    b = _d64(box).reshape(-1)
    v = np.zeros((1 << dim) * dim)
    lib().orc_box_vertices(dim, _ptr(b), _ptr(v))
    return v.reshape(1 << dim, dim)


def matrix_coeff_ref(dim, p):
    pp = _d64(p)
    a = np.zeros(3 if dim == 2 else 6)
    lib().orc_matrix_coeff_ref(dim, _ptr(pp), _ptr(a))
    return a


def fine_element(dim, x0, h, coef_kind, cdata, rhs_kind=RHS_CONST, rhs_value=1.0):
    nv = 1 << dim
    Ke = np.zeros(nv * nv)
    be = np.zeros(nv)
    lib().orc_fine_element(dim, _ptr(_d64(x0)), _ptr(_d64(h)), coef_kind, _ptr(_d64(cdata)), rhs_kind,
                           rhs_value, _ptr(Ke), _ptr(be))
    return Ke.reshape(nv, nv), be


def local_assemble(dim, nf, box, coef_kind, tile, rhs_kind=RHS_CONST, rhs_value=1.0):
    nn = n_nodes(dim, nf)
    K = np.zeros(nn * 3 ** dim)
    b = np.zeros(nn)
    lib().orc_local_assemble(dim, nf, _ptr(_d64(box).reshape(-1)), coef_kind, _ptr(_d64(tile).reshape(-1)),
                             rhs_kind, rhs_value, _ptr(K), _ptr(b))
    return K.reshape(nn, 3 ** dim), b


def stencil_apply(dim, nf, K, x):
    y = np.zeros(n_nodes(dim, nf))
    lib().orc_stencil_apply(dim, nf, _ptr(_d64(K).reshape(-1)), _ptr(_d64(x)), _ptr(y))
    return y


def stencil_to_dense(dim, nf, K):
    """Dense matrix of the structured stencil (small grids only)."""
    np1 = nf + 1
    nn = n_nodes(dim, nf)
    A = np.zeros((nn, nn))
    for n in range(nn):
        X, Y, Z = n % np1, (n // np1) % np1, (n // (np1 * np1) if dim == 3 else 0)
        for t in range(3 ** dim):
            ox, oy, oz = t % 3 - 1, (t // 3) % 3 - 1, (t // 9 - 1 if dim == 3 else 0)
            XX, YY, ZZ = X + ox, Y + oy, Z + oz
            if not (0 <= XX <= nf and 0 <= YY <= nf and 0 <= ZZ <= (nf if dim == 3 else 0)):
                continue
            A[n, XX + np1 * (YY + np1 * ZZ)] += K[n, t]
    return A


def dirichlet_data(dim, nf, box, index):
    g = np.zeros(n_nodes(dim, nf))
    lib().orc_dirichlet_data(dim, nf, _ptr(_d64(box).reshape(-1)), index, _ptr(g))
    return g


def local_bases_direct(dim, nf, box, K):
    phi = np.zeros((1 << dim) * n_nodes(dim, nf))
    rc = lib().orc_local_bases_direct(dim, nf, _ptr(_d64(box).reshape(-1)), _ptr(_d64(K).reshape(-1)), _ptr(phi))
    if rc:
        raise ValueError("local matrix not SPD")
    return phi.reshape(1 << dim, -1)


def local_bases_ssor_pcg(dim, nf, box, K, omega=1.6, abs_tol=1e-12, max_iter=1000):
    nb = 1 << dim
    phi = np.zeros(nb * n_nodes(dim, nf))
    its = np.zeros(nb, dtype=np.int32)
    fails = lib().orc_local_bases_ssor_pcg(dim, nf, _ptr(_d64(box).reshape(-1)), _ptr(_d64(K).reshape(-1)), omega,
                                           abs_tol, max_iter, _ptr(phi), _ptr(its, _ip))
    return phi.reshape(nb, -1), its, fails


def element(dim, nf, K, b, phi):
    nb = 1 << dim
    Ae = np.zeros(nb * nb)
    Fe = np.zeros(nb)
    lib().orc_element(dim, nf, _ptr(_d64(K).reshape(-1)), _ptr(_d64(b)), _ptr(_d64(phi).reshape(-1)), _ptr(Ae),
                      _ptr(Fe))
    return Ae.reshape(nb, nb), Fe


def batch_build(dim, box, nf, coef_kind, coef, rhs_kind=RHS_CONST, rhs_value=1.0, solver=SOLVER_DIRECT,
                omega=1.6, abs_tol=1e-12, max_iter=1000, n_threads=0, want_bases=False):
    """All cells of a batch.  box [n][dim][2], nf [n] (int), coef: concatenated tiles.
    Returns dict(Ae [n][nb][nb], Fe [n][nb], iters [n][nb], bases (flat or None), fails)."""
    box = _d64(box).reshape(-1)
    nf = np.ascontiguousarray(nf, dtype=np.int32)
    n = nf.shape[0]
    nb = 1 << dim
    Ae = np.zeros(n * nb * nb)
    Fe = np.zeros(n * nb)
    its = np.zeros(n * nb, dtype=np.int32)
    bases = None
    if want_bases:
        bases = np.zeros(int(sum(nb * (int(k) + 1) ** dim for k in nf)))
    fails = lib().orc_batch_build(dim, n, _ptr(box), _ptr(nf, _ip), coef_kind, _ptr(_d64(coef).reshape(-1)),
                                  rhs_kind, rhs_value, solver, omega, abs_tol, max_iter, n_threads, _ptr(bases),
                                  _ptr(Ae), _ptr(Fe), _ptr(its, _ip))
    return {"Ae": Ae.reshape(n, nb, nb), "Fe": Fe.reshape(n, nb), "iters": its.reshape(n, nb), "bases": bases,
            "fails": fails}


def max_threads():
    return lib().orc_max_threads()


def coarse_assemble(dim, nc, Ae, Fe, bc_kind):
    nn = (nc + 1) ** dim
    Kc = np.zeros(nn * 3 ** dim)
    Fc = np.zeros(nn)
    cm = np.zeros(nn, dtype=np.uint8)
    cv = np.zeros(nn)
    lib().orc_coarse_assemble(dim, nc, _ptr(_d64(Ae).reshape(-1)), _ptr(_d64(Fe).reshape(-1)), bc_kind, _ptr(Kc),
                              _ptr(Fc), _ptr(cm, _up), _ptr(cv))
    return Kc.reshape(nn, 3 ** dim), Fc, cm, cv


def coarse_solve(dim, nc, Kc, Fc, cmask, cval):
    u = np.zeros((nc + 1) ** dim)
    rc = lib().orc_coarse_solve(dim, nc, _ptr(_d64(Kc).reshape(-1)), _ptr(_d64(Fc)),
                                _ptr(np.ascontiguousarray(cmask, dtype=np.uint8), _up), _ptr(_d64(cval)), _ptr(u))
    if rc:
        raise ValueError("coarse matrix not SPD")
    return u


def reconstruct(dim, nf, phi, w):
    u = np.zeros(n_nodes(dim, nf))
    lib().orc_reconstruct(dim, nf, _ptr(_d64(phi).reshape(-1)), _ptr(_d64(w)), _ptr(u))
    return u


# ---- general quadrilateral cells (2D): vertices quad[4][2] in deal.II order ----
def quad_nodes(quad, nf):
    """Fine node coordinates [(nf+1)^2][2] of a general quadrilateral refined nf times."""
    q = _d64(quad).reshape(-1)
    out = np.zeros(((nf + 1) ** 2, 2))
    x = np.zeros(2)
    for n in range((nf + 1) ** 2):
        lib().orc_quad_node(_ptr(q), nf, n % (nf + 1), n // (nf + 1), _ptr(x))
        out[n] = x
    return out


def fine_element_quad(fv, coef_kind, cdata, rhs_kind=RHS_CONST, rhs_value=1.0):
    Ke = np.zeros(16)
    be = np.zeros(4)
    lib().orc_fine_element_quad(_ptr(_d64(fv).reshape(-1)), coef_kind, _ptr(_d64(cdata)), rhs_kind, rhs_value,
                                _ptr(Ke), _ptr(be))
    return Ke.reshape(4, 4), be


def local_assemble_quad(nf, quad, coef_kind, tile, rhs_kind=RHS_CONST, rhs_value=1.0):
    nn = n_nodes(2, nf)
    K = np.zeros(nn * 9)
    b = np.zeros(nn)
    lib().orc_local_assemble_quad(nf, _ptr(_d64(quad).reshape(-1)), coef_kind, _ptr(_d64(tile).reshape(-1)),
                                  rhs_kind, rhs_value, _ptr(K), _ptr(b))
    return K.reshape(nn, 9), b


def dirichlet_data_quad(nf, quad, index):
    g = np.zeros(n_nodes(2, nf))
    lib().orc_dirichlet_data_quad(nf, _ptr(_d64(quad).reshape(-1)), index, _ptr(g))
    return g


def local_bases_direct_quad(nf, quad, K):
    phi = np.zeros(4 * n_nodes(2, nf))
    rc = lib().orc_local_bases_direct_quad(nf, _ptr(_d64(quad).reshape(-1)), _ptr(_d64(K).reshape(-1)), _ptr(phi))
    if rc:
        raise ValueError("local matrix not SPD")
    return phi.reshape(4, -1)


def batch_build_quad(quad, nf, coef_kind, coef, rhs_kind=RHS_CONST, rhs_value=1.0, solver=SOLVER_DIRECT,
                     omega=1.6, abs_tol=1e-12, max_iter=1000, n_threads=0, want_bases=False):
    """batch_build for general quadrilaterals: quad [n][4][2]."""
    q = _d64(quad).reshape(-1)
    nf = np.ascontiguousarray(nf, dtype=np.int32)
    n = nf.shape[0]
    Ae = np.zeros(n * 16)
    Fe = np.zeros(n * 4)
    its = np.zeros(n * 4, dtype=np.int32)
    bases = np.zeros(int(sum(4 * (int(k) + 1) ** 2 for k in nf))) if want_bases else None
    fails = lib().orc_batch_build_quad(n, _ptr(q), _ptr(nf, _ip), coef_kind, _ptr(_d64(coef).reshape(-1)),
                                       rhs_kind, rhs_value, solver, omega, abs_tol, max_iter, n_threads,
                                       _ptr(bases), _ptr(Ae), _ptr(Fe), _ptr(its, _ip))
    return {"Ae": Ae.reshape(n, 4, 4), "Fe": Fe.reshape(n, 4), "iters": its.reshape(n, 4), "bases": bases,
            "fails": fails}


# ---- general hexahedral cells (3D): vertices hex[8][3] in deal.II order ----
def hex_nodes(hexa, nf):
    """Fine node coordinates [(nf+1)^3][3] of a general hexahedron refined nf times."""
    h = _d64(hexa).reshape(-1)
    np1 = nf + 1
    out = np.zeros((np1 ** 3, 3))
    x = np.zeros(3)
    I = np.zeros(3, np.int32)
    for n in range(np1 ** 3):
        I[:] = n % np1, (n // np1) % np1, n // (np1 * np1)
        lib().orc_cell_node(3, _ptr(h), nf, _ptr(I, _ip), _ptr(x))
        out[n] = x
    return out


def fine_element_hex(fv, coef_kind, cdata, rhs_kind=RHS_CONST, rhs_value=1.0):
    Ke = np.zeros(64)
    be = np.zeros(8)
    lib().orc_fine_element_hex(_ptr(_d64(fv).reshape(-1)), coef_kind, _ptr(_d64(cdata)), rhs_kind, rhs_value,
                               _ptr(Ke), _ptr(be))
    return Ke.reshape(8, 8), be


def local_assemble_hex(nf, hexa, coef_kind, tile, rhs_kind=RHS_CONST, rhs_value=1.0):
    nn = n_nodes(3, nf)
    K = np.zeros(nn * 27)
    b = np.zeros(nn)
    lib().orc_local_assemble_hex(nf, _ptr(_d64(hexa).reshape(-1)), coef_kind, _ptr(_d64(tile).reshape(-1)),
                                 rhs_kind, rhs_value, _ptr(K), _ptr(b))
    return K.reshape(nn, 27), b


def dirichlet_data_hex(nf, hexa, index):
    g = np.zeros(n_nodes(3, nf))
    lib().orc_dirichlet_data_hex(nf, _ptr(_d64(hexa).reshape(-1)), index, _ptr(g))
    return g


def local_bases_direct_hex(nf, hexa, K):
    phi = np.zeros(8 * n_nodes(3, nf))
    rc = lib().orc_local_bases_direct_hex(nf, _ptr(_d64(hexa).reshape(-1)), _ptr(_d64(K).reshape(-1)), _ptr(phi))
    if rc:
        raise ValueError("local matrix not SPD")
    return phi.reshape(8, -1)


def batch_build_hex(hexa, nf, coef_kind, coef, rhs_kind=RHS_CONST, rhs_value=1.0, solver=SOLVER_DIRECT,
                    omega=1.6, abs_tol=1e-12, max_iter=1000, n_threads=0, want_bases=False):
    """batch_build for general hexahedra: hexa [n][8][3]."""
    h = _d64(hexa).reshape(-1)
    nf = np.ascontiguousarray(nf, dtype=np.int32)
    n = nf.shape[0]
    Ae = np.zeros(n * 64)
    Fe = np.zeros(n * 8)
    its = np.zeros(n * 8, dtype=np.int32)
    bases = np.zeros(int(sum(8 * (int(k) + 1) ** 3 for k in nf))) if want_bases else None
    fails = lib().orc_batch_build_general(3, n, _ptr(h), _ptr(nf, _ip), coef_kind, _ptr(_d64(coef).reshape(-1)),
                                          rhs_kind, rhs_value, solver, omega, abs_tol, max_iter, n_threads,
                                          _ptr(bases), _ptr(Ae), _ptr(Fe), _ptr(its, _ip))
    return {"Ae": Ae.reshape(n, 8, 8), "Fe": Fe.reshape(n, 8), "iters": its.reshape(n, 8), "bases": bases,
            "fails": fails}


# ---------------------------------------------------------------------------
# A posteriori energy-norm bound of a local solve (no counterpart in the
# reference, whose only stopping rule is an absolute residual,
# include/base/diffusion_problem_basis.tpp:297).  For the interior equations
# K_II x = b of basis i with computed phi (boundary values exact) and interior
# residual r:
#   ||phi - phi*||_K^2 = r^T K_II^-1 r <= ||r||^2 / lam_lb,
#   lam_lb = min over the tile of lambda_min(A) * mu_min (Loewner: K(A) >= a K(I)),
#   ||phi*||_K^2 >= e_unit^2 / D_i  (complementary energy with the divergence-free
#   flux t grad N_i: N_i is multilinear, hence harmonic),
#   e_unit = int |grad N_i|^2,  D_i = int grad N_i^T A^-1 grad N_i.
# ---------------------------------------------------------------------------
def unit_stiffness_lambda_min(dim, nf, H):
    """Smallest eigenvalue of the unit-coefficient Q1 stiffness on the interior
    nodes of a box with side lengths H and nf fine cells per side.  The operator
    is a sum of tensor products of the 1D stiffness and mass, whose common
    eigenvectors are sin(j pi x); eigenvalue sum_a k_a prod_{b != a} m_b with
    k = (2 - 2 cos t)/h, m = h (4 + 2 cos t)/6, t = j pi / nf per axis.  The
    minimum over all j is taken (on strongly stretched boxes it need not sit at
    j = 1 on every axis)."""
    c = np.cos(np.arange(1, nf) * np.pi / nf)
    h = np.asarray(H, dtype=np.float64) / nf
    k = [(2 - 2 * c) / h[a] for a in range(dim)]
    m = [h[a] * (4 + 2 * c) / 6 for a in range(dim)]
    grids = np.meshgrid(*[np.arange(nf - 1)] * dim, indexing="ij")
    lam = sum(k[a][grids[a]] * np.prod([m[b][grids[b]] for b in range(dim) if b != a], axis=0) for a in range(dim))
    return float(lam.min())


def _cell_lambda_and_inverse(dim, kind, cdata):
    """(smallest eigenvalue over the cell's values, A^-1 to use in D) for one fine cell."""
    nq = 1 << dim
    if kind == COEF_SCALAR_CELL:
        a = float(cdata[0])
        return a, np.eye(dim) / a
    if kind == COEF_SCALAR_QP:
        a = float(np.min(cdata[:nq]))
        return a, np.eye(dim) / a
    ns = 3 if dim == 2 else 6
    idx = [(0, 0), (0, 1), (1, 1)] if dim == 2 else [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

    def full(v):
        A = np.zeros((dim, dim))
        for (i, j), x in zip(idx, v):
            A[i, j] = A[j, i] = x
        return A
    if kind == COEF_TENSOR_CELL:
        A = full(cdata[:ns])
        lam = float(np.linalg.eigvalsh(A)[0])
        # 3D: I / lambda_min (>= A^-1: a valid, simpler upper bound in D)
        return lam, (np.linalg.inv(A) if dim == 2 else np.eye(dim) / lam)
    lam = min(float(np.linalg.eigvalsh(full(cdata[q * ns:(q + 1) * ns]))[0]) for q in range(nq))
    return lam, np.eye(dim) / lam


def energy_bound_terms(dim, nf, box, coef_kind, tile):
    """-> (lam_lb, e_lb[nb]): the lower bounds above for one box cell."""
    box = _d64(box).reshape(dim, 2)
    H = box[:, 1] - box[:, 0]
    stride = coef_stride(dim, coef_kind)
    tile = _d64(tile).reshape(-1)
    nb = 1 << dim
    t = np.arange(nf + 1) / nf
    # per axis and fine index j: int over the cell of l^2 and of l, l = t or 1 - t (normalised)
    J0 = {1: (t[1:] ** 3 - t[:-1] ** 3) / 3, 0: ((1 - t[:-1]) ** 3 - (1 - t[1:]) ** 3) / 3}
    J1 = {1: (t[1:] ** 2 - t[:-1] ** 2) / 2, 0: ((1 - t[:-1]) ** 2 - (1 - t[1:]) ** 2) / 2}
    vol = float(np.prod(H))
    e_unit = vol * sum(1.0 / (3 ** (dim - 1) * H[k] ** 2) for k in range(dim))
    lam_min, D = np.inf, np.zeros(nb)
    for f in range(nf ** dim):
        j = [f % nf, (f // nf) % nf, f // (nf * nf)][:dim]
        lam, B = _cell_lambda_and_inverse(dim, coef_kind, tile[f * stride:(f + 1) * stride])
        lam_min = min(lam_min, lam)
        for i in range(nb):
            bit = [(i >> k) & 1 for k in range(dim)]
            s = [1.0 if b else -1.0 for b in bit]
            # G[k][m] = int_cell d_k N_i d_m N_i (exact for the multilinear N_i)
            G = np.zeros((dim, dim))
            for k in range(dim):
                for m in range(dim):
                    v = s[k] * s[m] / (H[k] * H[m])
                    for a in range(dim):
                        if a == k and a == m:
                            v *= H[a] / nf
                        elif a == k or a == m:
                            v *= H[a] * J1[bit[a]][j[a]]
                        else:
                            v *= H[a] * J0[bit[a]][j[a]]
                    G[k, m] = v
            D[i] += float(np.sum(B * G))
    lam_lb = lam_min * unit_stiffness_lambda_min(dim, nf, H)
    return lam_lb, e_unit ** 2 / D


def energy_bound(lam_lb, e_ref, eta):
    """Bound of ||phi - phi*||_K / ||phi*||_K from the residual norm eta and
    e_ref = ||phi||_K^2 of the computed phi (A^K_ii): ||phi*||^2 = ||phi||^2 - ||e||^2."""
    err2 = np.asarray(eta, dtype=np.float64) ** 2 / lam_lb
    den = np.asarray(e_ref, dtype=np.float64) - err2
    return np.where(den > 0, np.sqrt(err2 / np.maximum(den, 1e-300)), np.inf)
