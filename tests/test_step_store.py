"""The SQLite step records (CPU): the reference's flow and particle tables."""

from funlbm_amd.lbm.store import StepStore


def test_store_round_trip(tmp_path):
    s = StepStore(str(tmp_path / "run.db"))
    flow = {"f": [0.01, 0.05, 0.3], "u": [0.0, 1e-4, 2e-3], "rho": [0.99, 1.0, 1.01]}
    parts = [{"cx": [1.0, 2.0, 3.0]}, {"cx": [4.0, 5.0, 6.0]}]
    s.put(7, flow, parts)
    s.put(7, flow, parts)  # rewriting a step replaces it
    assert s.flow(7) == flow
    assert s.particle(7, 2) == parts[1]
    assert s.particle(7, 3) is None and s.flow(8) is None
    n = s.db.execute("SELECT COUNT(*) FROM particle").fetchone()[0]
    assert n == 2
    s.close()
