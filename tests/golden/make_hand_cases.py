"""Write tests/golden/hand_cases.json: tiny maps computed by hand, exactly.

Atoms are small dyadic rationals (exact in float32).  For every cell the
window is read off the time axis directly -- rect: atoms with
t0 <= t_i < t0 + tau -- and F = x^H M^-1 x is evaluated with exact rational
arithmetic (fractions.Fraction) as sum over the real and imaginary parts of
x^T M^-1 x, M = [[A, C], [C, B]], x = (Fa, Fb), with M^-1 taken by Gaussian
elimination rather than a closed form.  The guard cond < 1e4 is applied with
the exact eigenvalue ratio.  No reference or oracle code is run.

Run: python tests/golden/make_hand_cases.py
"""

import json
import os
from fractions import Fraction as Fr

TATOM = 1800
T0 = 1_000_000_000


def solve2(M, v):
    """Solve M y = v for 2x2 M by elimination with exact Fractions."""
    (a, b), (c, d) = M
    v0, v1 = v
    if a != 0:
        f = c / a
        d2 = d - f * b
        w1 = v1 - f * v0
        y1 = w1 / d2
        y0 = (v0 - b * y1) / a
    else:
        y0 = None
        y1 = v0 / b
        y0 = (v1 - d * y1) / c
    return y0, y1


def cond_ok(A, B, C):
    # cond = lmax/lmin < 1e4  <=>  (lmax + lmin)^2 / (lmax*lmin) < (1e4+1)^2/1e4
    D = A * B - C * C
    T = A + B
    if D <= 0 or T <= 0:
        return False
    return T * T * 10000 < (10001 ** 2) * D


def F_cell(atoms, idx):
    A = sum((atoms[i][0] for i in idx), Fr(0))
    B = sum((atoms[i][1] for i in idx), Fr(0))
    C = sum((atoms[i][2] for i in idx), Fr(0))
    far = sum((atoms[i][3] for i in idx), Fr(0))
    fai = sum((atoms[i][4] for i in idx), Fr(0))
    fbr = sum((atoms[i][5] for i in idx), Fr(0))
    fbi = sum((atoms[i][6] for i in idx), Fr(0))
    if not cond_ok(A, B, C):
        return Fr(2)
    M = ((A, C), (C, B))
    yr = solve2(M, (far, fbr))
    yi = solve2(M, (fai, fbi))
    return far * yr[0] + fbr * yr[1] + fai * yi[0] + fbi * yi[1]


def rect_case(name, atoms, ts, t0, t0Band, dt0, tau, tauBand, dtau):
    N_t0 = t0Band // dt0 + 1
    N_tau = tauBand // dtau + 1
    F = []
    for m in range(N_t0):
        row = []
        for n in range(N_tau):
            a = t0 + m * dt0
            b = a + tau + n * dtau
            idx = [i for i, t in enumerate(ts) if a <= t < b]
            row.append(float(F_cell(atoms, idx)))
        F.append(row)
    return {
        "name": name,
        "TAtom": TATOM,
        "timestamps": ts,
        "atoms": [[float(x) for x in a] for a in atoms],
        "window": [1, t0, t0Band, dt0, tau, tauBand, dtau],
        "F": F,
    }


def main():
    q = Fr(1, 8)
    atoms3 = [
        (Fr(9, 16), Fr(1, 16), Fr(3, 16), Fr(3, 4), -Fr(1, 2), Fr(1, 4), Fr(1, 8)),
        (Fr(1, 4), Fr(1, 4), -Fr(1, 8), -Fr(5, 8), Fr(3, 8), Fr(7, 8), -Fr(1, 4)),
        (Fr(1, 16), Fr(9, 16), Fr(1, 8), Fr(1, 2), Fr(1, 2), -Fr(3, 4), Fr(5, 8)),
    ]
    ts3 = [T0 + i * TATOM for i in range(3)]
    atoms5 = [
        (Fr(25, 64), Fr(1, 64), Fr(5, 64), q * 3, -q, q * 2, q * 5),
        (Fr(9, 64), Fr(4, 64), Fr(6, 64), -q * 4, q * 6, q, -q * 3),
        (Fr(1, 64), Fr(16, 64), -Fr(4, 64), q * 7, q * 2, -q * 6, q),
        (Fr(16, 64), Fr(9, 64), -Fr(12, 64), -q, -q * 5, q * 3, q * 4),
        (Fr(36, 64), Fr(1, 64), Fr(6, 64), q * 2, q * 2, -q * 2, -q * 7),
    ]
    # one gap: the fourth grid slot is empty
    ts5 = [T0 + i * TATOM for i in (0, 1, 2, 4, 5)]
    cases = [
        rect_case("three_atoms_full_grid", atoms3, ts3, T0, 1 * TATOM, TATOM, 2 * TATOM,
                  1 * TATOM, TATOM),
        rect_case("five_atoms_with_gap", atoms5, ts5, T0, 3 * TATOM, TATOM, 2 * TATOM,
                  3 * TATOM, TATOM),
    ]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hand_cases.json")
    with open(out, "w") as f:
        json.dump({"generator": "tests/golden/make_hand_cases.py", "cases": cases}, f,
                  indent=1)
    print(out)


if __name__ == "__main__":
    main()
