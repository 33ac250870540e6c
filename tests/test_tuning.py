"""Tuning-table bookkeeping (no GPU needed)."""
from clmsoil import tuning


def test_table_roundtrip_and_nearest_lookup(tmp_path, monkeypatch):
    p = tmp_path / "t.json"
    base = {"device": "gfx950", "workload": "production_day",
            "richards": {"g": 4, "block": 256, "persistent": True, "blocks_per_cu": 0},
            "heat": {"g": 4, "block": 256}}
    tuning.save_entry(dict(base, columns=1 << 20), p)
    tuning.save_entry(dict(base, columns=1 << 16, heat={"g": 2, "block": 128}), p)
    tuning.save_entry(dict(base, columns=1 << 20, heat={"g": 1, "block": 64}), p)   # replaces
    t = tuning.load_table(p)
    assert len(t) == 2
    monkeypatch.setattr(tuning, "TABLE", p)
    monkeypatch.setattr(tuning, "load_table", lambda path=p: tuning.json.loads(p.read_text()))
    e = tuning.lookup(900000, device="gfx950")
    assert e["columns"] == 1 << 20 and e["heat"] == {"g": 1, "block": 64}
    assert tuning.lookup(70000, device="gfx950")["columns"] == 1 << 16
    assert tuning.lookup(70000, device="gfx942") is None
