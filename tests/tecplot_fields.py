"""numpy restatement of the cell-centred output variables of Output/OUTPUT_InstField.F90:
168-370 (kind_field = 0), from a block's reference-layout state arrays, for the writer tests."""
import numpy as np

from trans_amd.mesh import O

NAMES = ("density", "U", "V", "W", "Uabs", "Vabs", "Wabs", "Ma", "Mabs", "Ps", "P0rel", "P0", "Ts", "T0rel", "T0",
         "Ox", "Oy", "Oz", "Om", "Q", "Yita4", "Rmiu", "k", "shock", "fr", "uu", "vv", "ww", "uv", "uw", "vw")


def field_vars(v, pp, t, rmiu, dq, shock, yc, zc, omega, scales=(1.0, 1.0, 1.0, 1.0)):
    """All inputs interior-only arrays (v: 7 x NI1 x NJ1 x NK1, dq: >= 9 x ...); returns
    {name: array} in the order of NAMES."""
    roref, vref, pref, tinf = scales
    den = v[0]
    uu, vv, ww = v[1] / den, v[2] / den, v[3] / den
    ua, va, wa = uu, vv - omega * zc, ww + omega * yc                       # :183-185
    a = np.sqrt(1.4 * pp / den)
    ma = np.sqrt(uu * uu + vv * vv + ww * ww) / a                           # :238
    mabs = np.sqrt(ua ** 2 + va ** 2 + wa ** 2)                              # :239, not over a
    crel = 1.0 + 0.2 * ma * ma
    cabs = 1.0 + mabs * mabs
    out = {"density": den * roref, "U": uu * vref, "V": vv * vref, "W": ww * vref,
           "Uabs": ua * vref, "Vabs": va * vref, "Wabs": wa * vref, "Ma": ma, "Mabs": mabs,
           "Ps": pp * pref, "P0rel": pp * crel ** 3.5 * pref, "P0": pp * (cabs * 3.5) * pref,  # :262-266
           "Ts": t * tinf, "T0rel": t * crel * tinf, "T0": t * cabs * tinf}
    g = dq[:9]
    grad = g.reshape((3, 3) + g.shape[1:])                                   # [velocity][direction]
    S = 0.5 * (grad + grad.transpose((1, 0) + tuple(range(2, grad.ndim))))
    W = 0.5 * (grad - grad.transpose((1, 0) + tuple(range(2, grad.ndim))))
    ox, oy, oz = W[2, 1], W[0, 2], W[1, 0]                                   # :310-312
    # Q and Yita4 summed in the reference's loop order (first index fastest, :318-333): Q is a
    # small difference of large sums, so the order decides its last digits
    qv = np.zeros_like(den)
    for e in range(3):
        for a_ in range(3):
            qv = qv + W[a_, e] * W[a_, e] - S[a_, e] * S[a_, e]
    y4 = np.zeros_like(den)
    for h in range(3):
        for e in range(3):
            for a_ in range(3):
                y4 = y4 + S[a_, e] * W[e, h] * W[h, a_]
    out.update(Ox=ox, Oy=oy, Oz=oz, Om=2.0 * np.sqrt(ox * ox + oy * oy + oz * oz), Q=0.5 * qv, Yita4=y4)
    rk = v[5]
    div = g[0] + g[4] + g[8]
    out.update(Rmiu=rmiu, k=rk / den, shock=shock, fr=np.ones_like(den),
               uu=(2 * g[0] - 2 * div / 3) * rmiu - 2 * rk / 3, vv=(2 * g[4] - 2 * div / 3) * rmiu - 2 * rk / 3,
               ww=(2 * g[8] - 2 * div / 3) * rmiu - 2 * rk / 3, uv=(g[1] + g[3]) * rmiu,
               uw=(g[2] + g[6]) * rmiu, vw=(g[5] + g[7]) * rmiu)
    return {n: out[n] for n in NAMES}


def interior_cells(a, b):
    """Interior cells 1..N-1 of a -2:N+2 reference-layout array (components first if 4-D)."""
    s = (slice(O + 1, O + b.ni), slice(O + 1, O + b.nj), slice(O + 1, O + b.nk))
    return a[(slice(None),) + s] if a.ndim == 4 else a[s]
