"""In-situ parity of the device AMEn site step against the oracle.

The oracle's ``amen`` hands a snapshot of its full state to a hook before
every site step (``oracle.tt_oracle.SiteState``: x, z and the ten
environments at that site).  :func:`device_site_step` loads one snapshot
into device memory and runs the device path's own site-step functions on it
(``ttnumpy_amd/solvers/amen.py``: local solve, residual core, SVD
truncation with the residual-budget rank rule, enrichment, the z shift), the
functions the per-kernel ``amen`` calls at ``ttnumpy/solvers/amen.py:579-708``.

Each stage runs twice: *fed*, on the oracle's own inputs to that stage (the
oracle's local solution for the residual core and the truncation, the
oracle's truncated core, carry and enrichment block for the enrichment), so
a stage is compared alone; and *chained*, on the device's own results from
the stage before, as the device solver runs.  :func:`compare` reduces both
to the quantities the parity tests bound.
"""

from __future__ import annotations

import numpy as np

from oracle import tt_oracle as O


def _dev():
    from ttnumpy_amd import device as D
    return D


def device_settings(d: int, residual_tol=1e-8, symmetric=True, direct_solve_size=1000,
                    max_sweeps=20):
    from ttnumpy_amd.solvers.common import _SolverSettings
    return _SolverSettings.resolve(
        max_sweeps=max_sweeps, residual_tol=residual_tol, stagnation_factor="auto",
        local_solver="auto", symmetric=symmetric, direct_solve_size=direct_solve_size,
        return_info=True, local_rtol=None, allow_dense_fallback=False,
        truncation_steps=max(2 * (d - 1), 1))


def _h(t):
    return t.detach().cpu().numpy()


def device_site_step(state: O.SiteState, A_d, b_d, settings, oracle_out: dict,
                     error_threshold=1e-7, max_rank=None, residual_damp=2.0) -> dict:
    """Run the device site step on ``state``; see the module docstring."""
    import importlib

    import ttnumpy_amd as tx
    from ttnumpy_amd.device import P
    # the module, not the function of the same name that the package exports
    AM = importlib.import_module("ttnumpy_amd.solvers.amen")
    from ttnumpy_amd.solvers.one_site import (_shift_center_left, _shift_center_right,
                                              _solve_local_one_site, local_rhs_one_site)
    D = _dev()
    k, d = state.k, len(state.x)
    env = {name: D.to_device(v) for name, v in state.envs.items()}

    def at(name):
        lst = [None] * d
        lst[k] = env[name]
        return lst

    def train(cores):
        t = tx.TensorTrain([np.asarray(c) for c in cores])
        return t

    mode = settings.truncation_mode
    tt = settings.truncation_tol
    out = {}
    x = train(state.x)
    b_loc = local_rhs_one_site(k, at("b_L"), at("b_R"), b_d)
    u, op, _ = _solve_local_one_site(k, x, A_d, b_d, at("A_L"), at("A_R"), at("b_L"), at("b_R"),
                                     settings, preserve_operator=True)
    out["b_loc"], out["u"] = _h(b_loc), _h(u)
    rl, n, rr = x.bonds[k], x.row_modes[k], x.bonds[k + 1]
    fed_core3 = D.to_device(oracle_out["u"].reshape(rl, n, rr))
    own_core3 = u.reshape(rl, n, rr)
    out["z_new_fed"] = _h(AM.residual_core(k, at("zb_L"), at("zb_R"), at("zA_L"), at("zA_R"), A_d,
                                           b_d, x, x_core=fed_core3))
    z_own = AM.residual_core(k, at("zb_L"), at("zb_R"), at("zA_L"), at("zA_R"), A_d, b_d, x,
                             x_core=own_core3)
    out["z_new"] = _h(z_own)
    if state.direction == "backward" and k == 0:
        return out
    fwd = state.direction == "forward"
    trunc = AM._truncate_core_right if fwd else AM._truncate_core_left
    # truncation on the oracle's local solution and b_loc
    spectra = []
    core_f, carry_f = trunc(fed_core3, mode, error_threshold, max_rank, op,
                            D.to_device(oracle_out["b_loc"]), tt, residual_damp, spectra=spectra)
    out["S_fed"], out["rank_fed"] = _h(spectra[0]), int(core_f.shape[2] if fwd else core_f.shape[0])
    # the whole step on the device's own results
    spectra = []
    core, carry = trunc(own_core3, mode, error_threshold, max_rank, op, b_loc, tt, residual_damp,
                        spectra=spectra)
    out["S"], out["rank"] = _h(spectra[0]), int(core.shape[2] if fwd else core.shape[0])

    def enrich(core, carry, block_src, z_src, xz):
        xt, zt = train(state.x), train(state.z)
        if fwd:
            xl = xz.shape[1]
            zl, nz, zr = z_src.shape
            block = D.empty(xl, nz, zr)
            D.gemm(1, xl, nz * zr, zl, 1.0, P(xz), 1, xl, 0, P(z_src), nz * zr, 1, 0, 0.0,
                   P(block), nz * zr, 1, 0)
            AM._enrich_right(xt, k, core, carry, block if block_src is None else block_src,
                             max_rank)
            _shift_center_right(zt, k, z_src)
            nb = k + 1
        else:
            block = D.absorb_right(z_src, P(xz), xz.shape[0], xz.shape[1])
            AM._enrich_left(xt, k, core, carry, block if block_src is None else block_src,
                            max_rank)
            _shift_center_left(zt, k, z_src)
            nb = k - 1
        return {"block": _h(block), "x_k": _h(xt.cores[k]), "x_nb": _h(xt.cores[nb]),
                "z_k": _h(zt.cores[k]), "z_nb": _h(zt.cores[nb])}

    xz = env["xz_L"] if fwd else env["xz_R"]
    fed = enrich(D.to_device(oracle_out["core"]), D.to_device(oracle_out["carry"]),
                 D.to_device(oracle_out["block"]), D.to_device(oracle_out["z_new"]), xz)
    out.update({k2 + "_fed": v for k2, v in fed.items()})
    own = enrich(core, carry, None, z_own, xz)
    out.update(own)
    return out


def _two_site(left, right, forward):
    """The product of the two cores a site step leaves at (k, neighbour):
    invariant under the gauge of the bond between them."""
    if forward:  # x_k (rl, n, 1, W), x_{k+1} (W, n', 1, rr')
        a = left.reshape(-1, left.shape[-1])
        return a @ right.reshape(right.shape[0], -1)
    a = right.reshape(-1, right.shape[-1])  # x_{k-1} (rl', n', 1, W)
    return a @ left.reshape(left.shape[0], -1)


def _frame(core, forward):
    """Orthonormal frame a QR shift leaves at k: left unfolding (forward) or
    the transposed right unfolding (backward)."""
    if forward:
        return core.reshape(-1, core.shape[-1])
    return core.reshape(core.shape[0], -1).T


def _proj_dist(Qa, Qb):
    """||Qa Qa^T - Qb Qb^T||_2 for orthonormal frames (the sine of the largest
    principal angle), or inf if the dimensions differ."""
    if Qa.shape != Qb.shape:
        return float("inf")
    # ||(I - Qa Qa^T) Qb||_2, accurate to rounding also for tiny angles
    return float(np.linalg.norm(Qb - Qa @ (Qa.T @ Qb), 2))


def margin(trace: dict) -> float:
    """Relative distance of the oracle's residual-budget decision from a flip:
    min over the partial sums it tested of |norm - thr| / thr (inf when the
    spectrum was zero or the threshold rule was used)."""
    parts = trace.get("partials")
    thr = trace.get("thr", 0.0)
    if not parts or thr <= 0.0:
        return float("inf")
    return float(min(abs(p - thr) for p in parts) / thr)


def compare(state: O.SiteState, o: dict, g: dict) -> dict:
    """Differences between the oracle's step ``o`` and the device's ``g``."""
    r = {"sweep": state.sweep, "dir": state.direction, "k": state.k, "N": o["u"].size}
    r["b_loc"] = float(np.linalg.norm(g["b_loc"] - o["b_loc"]) / np.linalg.norm(o["b_loc"]))
    r["u"] = float(np.linalg.norm(g["u"] - o["u"]) / np.linalg.norm(o["u"]))
    zo = o["z_new"].ravel()
    r["z_fed"] = float(np.linalg.norm(g["z_new_fed"].ravel() - zo) / o["z_scale"])
    r["z"] = float(np.linalg.norm(g["z_new"].ravel() - zo) / o["z_scale"])
    if "trace" not in o:
        return r
    fwd = state.direction == "forward"
    tr = o["trace"]
    So = tr["S"]
    r["sigma_fed"] = float(np.max(np.abs(g["S_fed"] - So)) / So[0]) if So[0] > 0 else 0.0
    r["sigma"] = float(np.max(np.abs(g["S"] - So)) / So[0]) if So[0] > 0 else 0.0
    r["rank_oracle"], r["rank_fed"], r["rank"] = tr["rank"], g["rank_fed"], g["rank"]
    r["margin"] = margin(tr)
    P_o = _two_site(o["x_k"], o["x_nb"], fwd)
    for sfx in ("_fed", ""):
        Pg = _two_site(g["x_k" + sfx], g["x_nb" + sfx], fwd)
        r["two_site" + sfx] = (float(np.linalg.norm(Pg - P_o) / np.linalg.norm(P_o))
                               if Pg.shape == P_o.shape else float("inf"))
        r["frame" + sfx] = _proj_dist(_frame(g["x_k" + sfx], fwd), _frame(o["x_k"], fwd))
        r["z_frame" + sfx] = _proj_dist(_frame(g["z_k" + sfx], fwd), _frame(o["z_k"], fwd))
    r["block_fed"] = float(np.linalg.norm(g["block_fed"] - o["block"]) /
                           max(np.linalg.norm(o["block"]), 1e-300))
    # The enriched core [core | block] that the QR shift factors is often
    # numerically rank deficient; its frame's columns beyond the numerical
    # rank are then an arbitrary orthonormal completion (any two QRs differ
    # there), so the frame is checked by what a QR guarantees instead: the
    # device frame Q is orthonormal and its span holds the oracle's enriched
    # core (E = Q Q^T E), and likewise for z's frame and the residual core.
    if fwd:
        enr = np.concatenate([o["core"], o["block"]], axis=2)
        enr = enr.reshape(-1, enr.shape[2])
    else:
        enr = np.concatenate([o["core"], o["block"]], axis=0)
        enr = enr.reshape(enr.shape[0], -1).T
    zo3 = o["z_new"]
    zmat = zo3.reshape(-1, zo3.shape[2]) if fwd else zo3.reshape(zo3.shape[0], -1).T
    s = np.linalg.svd(enr, compute_uv=False)
    r["enrich_cond"] = float(s[0] / s[-1]) if s[-1] > 0 else float("inf")
    for sfx in ("_fed", ""):
        Q = _frame(g["x_k" + sfx], fwd)
        Qz = _frame(g["z_k" + sfx], fwd)
        r["frame_orth" + sfx] = float(np.max(np.abs(Q.T @ Q - np.eye(Q.shape[1]))))
        r["z_frame_orth" + sfx] = float(np.max(np.abs(Qz.T @ Qz - np.eye(Qz.shape[1]))))
    Q = _frame(g["x_k_fed"], fwd)
    Qz = _frame(g["z_k_fed"], fwd)
    r["frame_span_fed"] = float(np.linalg.norm(enr - Q @ (Q.T @ enr)) / np.linalg.norm(enr))
    zn = max(np.linalg.norm(zmat), 1e-300)
    r["z_frame_span_fed"] = float(np.linalg.norm(zmat - Qz @ (Qz.T @ zmat)) / zn)
    return r
