"""Snapshot records and dataset files in the reference's formats (host side:
no GPU needed)."""
import numpy as np
import pytest

# trace_viewer/snapshot.py:14-98, in declaration order
SNAPSHOT_FIELDS = ["u", "v", "p", "mask", "dt", "nu", "dx", "dy", "divergence", "Re", "CFL",
                   "flow_type", "solver_metadata", "iteration", "timestamp"]


def _snap(nx=12, ny=8, **kw):
    from aerojax_amd.snapshot import Snapshot
    rng = np.random.default_rng(0)
    f = lambda: rng.normal(size=(nx, ny))  # noqa: E731
    return Snapshot(u=f(), v=f(), p=f(), mask=np.ones((nx, ny)), dt=1e-3, nu=5e-4, dx=0.1,
                    dy=0.05, **kw)


def test_snapshot_fields_match_reference_record():
    s = _snap()
    assert list(s.to_dict()) == SNAPSHOT_FIELDS
    assert (s.nx, s.ny, s.shape) == (12, 8, (12, 8))
    with pytest.raises(ValueError):
        _snap(divergence=np.zeros((3, 3)))


def test_snapshot_npz_roundtrip(tmp_path):
    from aerojax_amd.snapshot import Snapshot
    s = _snap(divergence=np.full((12, 8), 0.5), Re=1234.5, CFL=0.25, flow_type="von_karman",
              solver_metadata={"mg": "textbook", "cycles": [8, 8]}, iteration=7, timestamp=1.5)
    path = str(tmp_path / "snap.npz")
    s.save(path)
    r = Snapshot.load(path)
    for k, v in s.to_dict().items():
        w = getattr(r, k)
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, w), k
        else:
            assert v == w, k
    # None fields are omitted and come back as None
    s2 = _snap()
    s2.save(path)
    r2 = Snapshot.load(path)
    assert r2.divergence is None and r2.Re is None and r2.solver_metadata is None
    # the file holds no pickled objects
    with np.load(path, allow_pickle=False) as z:
        assert all(z[k].dtype != object for k in z.files)


def test_load_dataset_reference_convention(tmp_path):
    """A file in the reference writer's layout (training_utils.py:35-51) loads
    without unpickling; missing keys are reported."""
    from aerojax_amd.snapshot import DATASET_KEYS, load_dataset
    n, nx, ny = 3, 16, 8
    rng = np.random.default_rng(1)
    data = {k: rng.normal(size=(n, nx, ny)) for k in ("u", "v", "p", "mask", "divergence")}
    data.update(dx=0.25, dy=0.125, dt=1e-3, nx=nx, ny=ny)
    assert set(data) == set(DATASET_KEYS)
    path = str(tmp_path / "training_data.npz")
    np.savez(path, **data)
    d = load_dataset(path)
    assert d["nx"] == nx and d["dt"] == 1e-3 and d["p"].shape == (n, nx, ny)
    np.savez(path, **{k: v for k, v in data.items() if k != "p"})
    with pytest.raises(KeyError):
        load_dataset(path)
