"""Station geometry and file formats (ccfj.io): pair order, great-circle
distances, day-file gathering, summed.npz."""
import numpy as np

from ccfj import GetStationPairs
from ccfj import io


def test_pair_index_matches_station_pair_order():
    nsta = 7
    p = GetStationPairs(nsta).reshape(-1, 2)
    assert np.array_equal(io.pair_index(p[:, 0], p[:, 1], nsta), np.arange(len(p)))


def test_great_circle_closed_forms():
    R = io.EARTH_RADIUS_KM
    assert abs(io.great_circle_km(0, 0, 0, 1) - 2 * np.pi * R / 360) < 1e-9
    assert abs(io.great_circle_km(90, 0, 0, 123) - np.pi * R / 2) < 1e-9
    assert abs(io.great_circle_km(10, 20, -10, 200) - np.pi * R) < 1e-6
    rng = np.random.default_rng(0)
    la1, lo1, la2, lo2 = rng.uniform(-80, 80, 4 * 50).reshape(4, 50)
    d = io.great_circle_km(la1, lo1, la2, lo2)
    assert np.allclose(d, io.great_circle_km(la2, lo2, la1, lo1))
    # haversine on the same sphere (well conditioned away from antipodes)
    p1, p2 = np.radians(la1), np.radians(la2)
    h = np.sin((p2 - p1) / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(np.radians(lo2 - lo1) / 2) ** 2
    assert np.allclose(d, 2 * R * np.arcsin(np.sqrt(h)), rtol=1e-10)


def test_stalist_and_pair_distances(tmp_path):
    path = tmp_path / "stalist"
    path.write_text("A -91.0 38.0\nB -95.0 43.0\n\nC -91.5 43.8\n")
    names, lon, lat = io.read_stalist(path)
    assert list(names) == ["A", "B", "C"]
    r = io.pair_distances(lon, lat, GetStationPairs(3))
    assert np.allclose(r, [io.great_circle_km(38, -91, 43, -95), io.great_circle_km(38, -91, 43.8, -91.5),
                           io.great_circle_km(43, -95, 43.8, -91.5)])
    assert np.allclose(io.pair_distances([0, 3], [0, 4], [0, 1], geographic=False), [5.0])


def test_gather_days_and_summed(tmp_path):
    rng = np.random.default_rng(1)
    stalist = ["S0", "S1", "S2", "S3"]
    nf = 5
    f = np.arange(nf) * 0.1
    # station spectra of two days; day 2 lists its stations out of order
    days = [["S0", "S1", "S3"], ["S3", "S2", "S0"]]
    want = np.zeros((6, nf), complex)
    cnt = np.zeros(6)
    paths = []
    for k, names in enumerate(days):
        spec = {n: rng.normal(size=nf) + 1j * rng.normal(size=nf) for n in names}
        pairs = GetStationPairs(len(names)).reshape(-1, 2)
        ncfs = np.array([np.conj(spec[names[a]]) * spec[names[b]] for a, b in pairs], np.complex64)
        path = tmp_path / f"day{k}.npz"
        np.savez(path, ncfs=ncfs, StationPairs=pairs.reshape(-1), stalist=np.array(names), f=f)
        paths.append(path)
        for a, b in pairs:
            i, j = sorted((stalist.index(names[a]), stalist.index(names[b])))
            want[io.pair_index(i, j, 4)] += np.conj(spec[stalist[i]]) * spec[stalist[j]]
            cnt[io.pair_index(i, j, 4)] += 1
    got, count, ff = io.gather_days(paths, stalist)
    assert np.array_equal(count, cnt)
    has = cnt > 0
    assert np.allclose(got[has], (want[has] / cnt[has, None]).astype(np.complex64), atol=1e-6)
    assert np.all(got[~has] == 0)
    r_km = np.array([30.0, 10.0, 50.0, 20.0, 40.0, 60.0])
    io.save_summed(tmp_path / "summed.npz", got, r_km, ff)
    ncfs, r_m, f2 = io.load_summed(tmp_path / "summed.npz")
    assert np.array_equal(r_m, np.sort(r_km) * 1e3)
    assert np.array_equal(ncfs, got[np.argsort(r_km)])
    assert np.array_equal(f2, f)
