"""FP64 numpy reference of exponential-window candidates on a fixed support --
TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by
the product).

A candidate (t0, tau) of one template with support bins k_lo..k_hi has window
sums S_c = sum_k w_k^{p_c} x_{c,k}, w_k = exp(-(t_k - t0)/tau), p = (2,2,2,1,1,1,1)
(pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:82-102: win for Fa/Fb,
win^2 for a2/b2/ab) and F from them with the cond < 1e4 guard
(cudaTransientFstatExpWindow.cu:109-129).  Every weight here is a direct
``np.exp``; the derivatives are the closed forms of F(S) and S(t0, tau).

Pinned by tests/test_refine.py: at a grid cell with the map's support, F equals
the C oracle's map cell (oracle/tcw_oracle.c, direct sum); the analytic
derivatives equal FP64 central finite differences of ``evaluate``; dF/dt0
vanishes and F is invariant under t0 shifts on a fixed support (closed form);
``support_bruteforce`` is a per-bin loop over the map kernel's index and
in-window tests.
"""

import numpy as np

P = np.array([2, 2, 2, 1, 1, 1, 1])
COND_MAX = 1e4


def support_bruteforce(t0, tau, t_data, N, TA):
    """Bins with non-zero weight in the map's exponential window at an integer
    cell (t0, tau): i_t0 <= k <= i_t1 and t0 <= t_k <= t0 + 3 tau, with the
    uint32 index maths of cudaTransientFstatExpWindow.cu:27-45 (no wrap-around
    for the inputs used here).  Returns (k_lo, k_hi) or None if empty."""
    half = TA // 2
    t1 = t0 + 3 * tau
    i_t0 = max(0, min((t0 - t_data + half) // TA, N - 1))
    i_t1 = max(0, min((t1 - t_data + half) // TA - 1, N - 1))
    ks = [k for k in range(i_t0, i_t1 + 1) if t0 <= t_data + k * TA <= t1]
    if not ks:
        return None
    return ks[0], ks[-1]


def window_sums(cols, t_data, TA, k_lo, k_hi, t0, tau):
    """(S[7], U[7]) of one candidate: S as above, U_c = sum w^p (t_k - t0) x."""
    k = np.arange(k_lo, k_hi + 1)
    u = (t_data + k.astype(np.float64) * TA) - t0
    w = np.exp(-u / tau)
    x = cols[:, k].astype(np.float64)
    wp = np.where(P[:, None] == 2, w * w, w)
    S = (wp * x).sum(axis=1)
    U = (wp * u * x).sum(axis=1)
    return S, U


def F_of_sums(S):
    """F with the cond guard (FP64, condition number from the eigenvalues of
    [[A, C], [C, B]] as in cudaTransientFstatExpWindow.cu:109-129)."""
    A, B, C, far, fai, fbr, fbi = S
    s = A + B
    disc = np.sqrt((A - B) ** 2 + 4 * C * C)
    den = s - disc
    cond = (s + disc) / den if den > 0 else np.inf
    if not cond < COND_MAX:
        return 2.0, False
    D = A * B - C * C
    F = (B * (far * far + fai * fai) + A * (fbr * fbr + fbi * fbi)
         - 2 * C * (far * fbr + fai * fbi)) / D
    return F, True


def evaluate(cols, t_data, TA, k_lo, k_hi, t0, tau):
    """F, dF/dt0, dF/dtau and dF/dS (7) of one candidate."""
    S, U = window_sums(cols, t_data, TA, k_lo, k_hi, t0, tau)
    F, ok = F_of_sums(S)
    if not ok:
        return F, 0.0, 0.0, np.zeros(7)
    A, B, C, far, fai, fbr, fbi = S
    D = A * B - C * C
    p = far * far + fai * fai
    q = fbr * fbr + fbi * fbi
    r = far * fbr + fai * fbi
    g = np.array([
        (q - F * B) / D,
        (p - F * A) / D,
        2 * (F * C - r) / D,
        2 * (B * far - C * fbr) / D,
        2 * (B * fai - C * fbi) / D,
        2 * (A * fbr - C * far) / D,
        2 * (A * fbi - C * fai) / D,
    ])
    dt0 = float((P * S * g).sum() / tau)
    dtau = float((P * U * g).sum() / tau ** 2)
    return F, dt0, dtau, g


def atom_grad(cols_shape, t_data, TA, k_lo, k_hi, t0, tau, g, gF=1.0):
    """Gradient of gF*F with respect to the merged atoms [7, n] of one
    candidate's template (zero outside the support)."""
    out = np.zeros(cols_shape)
    k = np.arange(k_lo, k_hi + 1)
    w = np.exp(-((t_data + k.astype(np.float64) * TA) - t0) / tau)
    for c in range(7):
        out[c, k] = gF * g[c] * (w * w if P[c] == 2 else w)
    return out


def batch_evaluate(cols, m_off, t_data, TA, cand_tmpl, k_lo, k_hi, t0, tau):
    """evaluate() over candidates of many templates (merged layout of
    tcw_merge_batch).  Returns F [n], dt0 [n], dtau [n], g [n, 7]."""
    n = len(cand_tmpl)
    F = np.zeros(n)
    dt0 = np.zeros(n)
    dtau = np.zeros(n)
    g = np.zeros((n, 7))
    for i in range(n):
        t = int(cand_tmpl[i])
        sub = cols[:, m_off[t]:m_off[t + 1]]
        F[i], dt0[i], dtau[i], g[i] = evaluate(sub, float(t_data[t]), TA, int(k_lo[i]),
                                               int(k_hi[i]), float(t0[i]), float(tau[i]))
    return F, dt0, dtau, g


def scan_max_tau(cols, t_data, TA, k_lo, k_hi, t0, tau_lo, tau_hi, n=2001):
    """Largest F over a uniform grid of n taus in [tau_lo, tau_hi] (fixed
    support), and its tau."""
    taus = np.linspace(tau_lo, tau_hi, n)
    Fs = np.array([F_of_sums(window_sums(cols, t_data, TA, k_lo, k_hi, t0, x)[0])[0]
                   for x in taus])
    j = int(np.argmax(Fs))
    return Fs[j], taus[j]
