"""Write matvis_amd/data/astrometry_2015_2045.npz: daily samples of the
date-dependent quantities of ERFA's astrometry context (what eraApco13
derives from the date alone), for the host astrometry model to interpolate.

Run with an interpreter that has pyerfa, e.g.
``/opt/conda/bin/python3.9 matvis_amd/data/make_astrometry_table.py``.

At 0h UTC of every day from 2015-01-01 to 2045-01-01 (TT from ERFA's own
leap-second table; TDB taken as TT):

* ``x``, ``y``: the CIP coordinates of the IAU 2006/2000A
  bias-precession-nutation matrix (eraPnm06a, eraBpn2xy);
* ``s``: the CIO locator (eraS06);
* ``eh``: the Earth's heliocentric position (au) and ``vb`` its
  barycentric velocity (au/day), from eraEpv00.

These are the inputs eraApco13 passes to eraApco; the site-dependent rest
(Earth rotation angle, the observer's position and velocity) is computed at
run time.  Only pyerfa is used.
"""
from pathlib import Path

import erfa
import numpy as np

START, END = 2457023.5, 2467980.5  # 2015-01-01, 2045-01-01 (UTC JD)


def main():
    jd = np.arange(START, END + 0.5, 1.0)
    tai1, tai2 = erfa.utctai(jd, np.zeros_like(jd))
    tt1, tt2 = erfa.taitt(tai1, tai2)
    x, y = erfa.bpn2xy(erfa.pnm06a(tt1, tt2))
    s = erfa.s06(tt1, tt2, x, y)
    pvh, pvb = erfa.epv00(tt1, tt2)
    eh = np.stack([pvh["p"][:, i] for i in range(3)], axis=1)
    vb = np.stack([pvb["v"][:, i] for i in range(3)], axis=1)
    out = Path(__file__).with_name("astrometry_2015_2045.npz")
    np.savez_compressed(out, jd_utc0=np.float64(START), x=x, y=y, s=s,
                        eh=eh, vb=vb, erfa_version=erfa.__version__)
    print(f"wrote {out} ({jd.size} days)")


if __name__ == "__main__":
    main()
