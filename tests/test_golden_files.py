"""The committed golden records are well formed (CPU)."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def test_rmat22_golden_record_shape():
    gold = json.load(open(os.path.join(HERE, "golden", "rmat22_k256_golden.json")))
    assert gold["n"] == 4194304 and gold["m"] > 120_000_000
    assert len(gold["levels"]) == 201 and all(len(r) == 6 for r in gold["levels"])
    assert gold["levels"][-1][2] == gold["n"] and gold["cut"] * 2 == gold["levels"][-1][0]
    assert len(gold["part_sha256"]) == 64
