"""Per-step records in SQLite, as the reference's run loop keeps them
(src/funlbm/lbm/base.py:78-82,108-110): a key-value table `flow` (step ->
the flow's JSON statistics) and a two-key table `particle` ((step, particle
number from 1) -> the particle's JSON).  The reference uses an external
key-value store whose schema is its own; these two plain tables hold the same
records with the standard library's sqlite3."""

import json
import sqlite3


class StepStore:
    def __init__(self, path: str):
        self.path = path
        self.db = sqlite3.connect(path)
        self.db.execute("CREATE TABLE IF NOT EXISTS flow (key TEXT PRIMARY KEY, value TEXT)")
        self.db.execute(
            "CREATE TABLE IF NOT EXISTS particle (key1 TEXT, key2 TEXT, value TEXT, PRIMARY KEY (key1, key2))"
        )
        self.db.commit()

    def put(self, step: int, flow: dict, particles) -> None:
        self.db.execute("INSERT OR REPLACE INTO flow VALUES (?, ?)", (str(step), json.dumps(flow)))
        self.db.executemany(
            "INSERT OR REPLACE INTO particle VALUES (?, ?, ?)",
            [(str(step), str(i + 1), json.dumps(p)) for i, p in enumerate(particles)],
        )
        self.db.commit()

    def flow(self, step: int):
        row = self.db.execute("SELECT value FROM flow WHERE key = ?", (str(step),)).fetchone()
        return None if row is None else json.loads(row[0])

    def particle(self, step: int, number: int):
        row = self.db.execute(
            "SELECT value FROM particle WHERE key1 = ? AND key2 = ?", (str(step), str(number))
        ).fetchone()
        return None if row is None else json.loads(row[0])

    def close(self) -> None:
        self.db.close()
