"""Station geometry and on-disk formats around the F-J transform.

The reference's examples (examples/example_CC.ipynb cells 3-5,
examples/example_noise.py:6-11) keep their data in these forms; the helpers
here read and write them with numpy only (no pickles, no geopy):

  station list   text, one station per line: "name lon lat" (degrees)
                 (examples/stalist)
  day file       npz {ncfs complex64 [npairs_day, nf], StationPairs int32
                 [2 npairs_day], stalist [nsta_day], f [nf]}, the pairs of
                 that day's stations in GetStationPairs order
  summed.npz     npz {ncfs complex64 [npairs, nf] sorted by r, r [npairs]
                 in km, f [nf]}: the day files averaged per pair over the
                 days that hold it, distances great-circle

Pair (i, j), i < j, of nsta stations sits at index
i (2 nsta - i - 1) / 2 + j - i - 1 of the GetStationPairs order.
"""
import numpy as np

EARTH_RADIUS_KM = 6371.009   # mean Earth radius (IUGG)


def read_stalist(path):
    """Station list -> (names, lon, lat) (arrays of str, float64, float64)."""
    names, lon, lat = [], [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 3:
                continue
            names.append(parts[0])
            lon.append(float(parts[1]))
            lat.append(float(parts[2]))
    return np.array(names), np.array(lon), np.array(lat)


def pair_index(i, j, nsta):
    """Index of pair (i, j), i < j, in the GetStationPairs order."""
    i, j = np.asarray(i), np.asarray(j)
    return i * (2 * nsta - i - 1) // 2 + j - i - 1


def great_circle_km(lat1, lon1, lat2, lon2, radius=EARTH_RADIUS_KM):
    """Great-circle distance on a sphere (degrees in, km out), by the
    arctangent form of the spherical Vincenty formula, which stays accurate
    for both tiny and near-antipodal separations."""
    p1, p2 = np.radians(lat1), np.radians(lat2)
    dl = np.radians(np.asarray(lon2) - np.asarray(lon1))
    s1, c1, s2, c2 = np.sin(p1), np.cos(p1), np.sin(p2), np.cos(p2)
    sd, cd = np.sin(dl), np.cos(dl)
    num = np.hypot(c2 * sd, c1 * s2 - s1 * c2 * cd)
    den = s1 * s2 + c1 * c2 * cd
    return radius * np.arctan2(num, den)


def pair_distances(lon, lat, pairs, geographic=True):
    """Distances of the pairs (flattened GetStationPairs array): great
    circle in km for geographic coordinates (degrees), else Euclidean in
    the coordinates' unit."""
    p = np.asarray(pairs).reshape(-1, 2)
    a, b = p[:, 0], p[:, 1]
    lon, lat = np.asarray(lon, np.float64), np.asarray(lat, np.float64)
    if geographic:
        return great_circle_km(lat[a], lon[a], lat[b], lon[b])
    return np.hypot(lon[a] - lon[b], lat[a] - lat[b])


def gather_days(paths, stalist):
    """Average day files over the days that hold each pair of the full
    station list.  Returns (ncfs complex64 [npairs, nf], count [npairs],
    f)."""
    stalist = list(stalist)
    nsta = len(stalist)
    npairs = nsta * (nsta - 1) // 2
    acc, count, f = None, np.zeros(npairs), None
    for path in paths:
        with np.load(path, allow_pickle=False) as d:
            idx = np.array([stalist.index(s) for s in d["stalist"]])
            ii, jj = np.triu_indices(len(idx), 1)
            a, b = idx[ii], idx[jj]
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            # a day file's pair (ii, jj) is stored as conj(A_ii) B_jj; when the
            # global order swaps the stations, the product is the conjugate
            rows = d["ncfs"]
            rows = np.where((a > b)[:, None], np.conj(rows), rows)
            gid = pair_index(lo, hi, nsta)
            if acc is None:
                acc = np.zeros((npairs, rows.shape[1]), np.complex128)
                f = d["f"]
            np.add.at(acc, gid, rows)
            np.add.at(count, gid, 1)
    out = np.zeros(acc.shape, np.complex64)
    has = count > 0
    out[has] = (acc[has] / count[has, None]).astype(np.complex64)
    return out, count, f


def save_summed(path, ncfs, r_km, f):
    """Sort pairs by distance and write summed.npz {ncfs, r, f}."""
    idx = np.argsort(r_km)
    np.savez(path, ncfs=np.asarray(ncfs, np.complex64)[idx], r=np.asarray(r_km)[idx],
             f=np.asarray(f))


def load_summed(path):
    """summed.npz -> (ncfs complex64 [npairs, nf], r in metres, f); the
    form fj_noise takes (examples/example_noise.py:6-11 scale r by 1e3)."""
    with np.load(path, allow_pickle=False) as d:
        return d["ncfs"], d["r"] * 1e3, d["f"]
