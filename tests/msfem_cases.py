"""Test-side helpers over the package's workload definitions: the oracle's build
of a case and the energy-norm error against it."""
import numpy as np

import msfem
from workloads import *  # noqa: F401,F403  (the workload generators)
from workloads import split_bases


def oracle_build(case, want_bases=True, solver=0, n_threads=0):
    import oracle as O
    if case.get("vertices") is not None:
        build = O.batch_build_quad if case["dim"] == 2 else O.batch_build_hex
        return build(case["vertices"], case["n_fine"], case["coef_kind"], case["coef"], case["rhs_kind"],
                     case["rhs_value"], solver=solver, n_threads=n_threads, want_bases=want_bases)
    return O.batch_build(case["dim"], case["box"], case["n_fine"], case["coef_kind"], case["coef"], case["rhs_kind"],
                         case["rhs_value"], solver=solver, n_threads=n_threads, want_bases=want_bases)


def energy_error(case, gpu_flat, ref_flat):
    """max over bases of ||phi - phi_ref||_K / ||phi_ref||_K with the cell's fine stiffness K."""
    import oracle as O
    dim = case["dim"]
    g = split_bases(case, gpu_flat)
    r = split_bases(case, ref_flat)
    stride = msfem.coef_stride(dim, case["coef_kind"])
    off, worst = 0, 0.0
    for c, nf in enumerate(case["n_fine"]):
        nf = int(nf)
        n = nf ** dim * stride
        if case.get("vertices") is not None:
            assemble = O.local_assemble_quad if dim == 2 else O.local_assemble_hex
            K, _ = assemble(nf, case["vertices"][c], case["coef_kind"], case["coef"][off:off + n], case["rhs_kind"],
                            case["rhs_value"])
        else:
            K, _ = O.local_assemble(dim, nf, case["box"][c], case["coef_kind"], case["coef"][off:off + n],
                                    case["rhs_kind"], case["rhs_value"])
        off += n
        for i in range(1 << dim):
            d = g[c][i] - r[c][i]
            e2 = d @ O.stencil_apply(dim, nf, K, d)
            n2 = r[c][i] @ O.stencil_apply(dim, nf, K, r[c][i])
            worst = max(worst, np.sqrt(max(e2, 0.0) / n2))
    return worst
