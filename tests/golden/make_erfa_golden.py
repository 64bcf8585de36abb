"""Write tests/golden/erfa_vectors.npz: ERFA's own astrometry at the array
site, from pyerfa (run with an interpreter that has pyerfa, e.g.
``/opt/conda/bin/python3.9 tests/golden/make_erfa_golden.py``).

For a few UTC times at the HERA site (longitude 21.4283 E, latitude
-30.7215, 1073 m; UT1 - UTC = 0, no polar motion, no refraction):

* ``ctx_*``: the astrometry context of ``eraApco13`` -- heliocentric
  direction ``eh`` and distance ``em`` of the observer, its barycentric
  velocity ``v`` (units of c) and ``bm1 = sqrt(1 - |v|^2)``, the
  bias-precession-nutation matrix ``bpn``, the local Earth rotation angle
  ``eral``, polar-motion terms ``xpl``, ``ypl``, ``sin/cos`` of the latitude
  and the diurnal aberration factor;
* for 48 sources (RA uniform, sin(dec) uniform, seeded, plus the four
  cardinal points): ``ld`` = eraLd(1, p, p, eh, em, 1e-6) of the ICRS
  direction p (light deflection by the Sun, the limiter the reference
  passes); ``ab`` = eraAb(ld, v, em, bm1) (aberration);
* ``obs_enu``: (east, north, up) unit vectors of eraAtco13's observed
  azimuth and zenith distance (the full ERFA chain, diurnal aberration
  included, no refraction).

Only pyerfa is used; nothing from matvis or this repository.
"""
from pathlib import Path

import erfa
import numpy as np

LON, LAT, HM = np.radians(21.4283038269), np.radians(-30.7215261207), 1073.0
UTC = [2459863.0, 2459863.1375, 2459863.4, 2459953.25, 2460100.8]


def main():
    rng = np.random.default_rng(2024)
    ra = np.concatenate([rng.uniform(0, 2 * np.pi, 44),
                         [0.0, np.pi / 2, np.pi, 0.3]])
    sdec = rng.uniform(-1, 1, 44)
    dec = np.concatenate([np.arcsin(sdec), [0.0, 0.0, 0.5, -np.pi / 2 + 1e-3]])
    p = np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra),
                  np.sin(dec)], axis=-1)  # (nsrc, 3)
    out = {"ra": ra, "dec": dec, "utc": np.array(UTC),
           "site": np.array([LON, LAT, HM])}
    keys = ["eh", "em", "v", "bm1", "bpn", "eral", "xpl", "ypl", "sphi",
            "cphi", "diurab", "along"]
    ctx = {k: [] for k in keys}
    lds, abs_, enus = [], [], []
    for jd in UTC:
        astrom, eo = erfa.apco13(jd, 0.0, 0.0, LON, LAT, HM, 0.0, 0.0, 0.0,
                                 0.0, 0.0, 0.0)
        for k in keys:
            ctx[k].append(np.asarray(astrom[k], dtype=np.float64))
        eh = np.asarray(astrom["eh"], dtype=np.float64)
        em = float(astrom["em"])
        v = np.asarray(astrom["v"], dtype=np.float64)
        bm1 = float(astrom["bm1"])
        ld = erfa.ld(1.0, p, p, eh, em, 1e-6)
        ab = erfa.ab(ld, v, em, bm1)
        lds.append(ld)
        abs_.append(ab)
        aob, zob, hob, dob, rob, _ = erfa.atco13(
            ra, dec, 0.0, 0.0, 0.0, 0.0, jd, 0.0, 0.0, LON, LAT, HM, 0.0, 0.0,
            0.0, 0.0, 0.0, 0.0)
        enus.append(np.stack([np.sin(zob) * np.sin(aob),
                              np.sin(zob) * np.cos(aob), np.cos(zob)]))
    for k in keys:
        out["ctx_" + k] = np.array(ctx[k])
    out["ld"] = np.array(lds)  # (ntimes, nsrc, 3)
    out["ab"] = np.array(abs_)
    out["obs_enu"] = np.array(enus)  # (ntimes, 3, nsrc)
    path = Path(__file__).with_name("erfa_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
