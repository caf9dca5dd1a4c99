"""Copies the reference's saved throughput-network configuration (a JSON data
file, neural_lod/scenes/arcade/learned_weights/arcade_world_final_512_throughput_config.json)
into tests/golden/ as a fixture for the configuration parser.  Run once in the
build container:  python tests/golden/make_configs.py
"""
import json
from pathlib import Path

SRC = Path("/root/reference/neural_lod/scenes/arcade/learned_weights/arcade_world_final_512_throughput_config.json")
DST = Path(__file__).resolve().parent / "reference_throughput_config.json"

if __name__ == "__main__":
    cfg = json.loads(SRC.read_text())
    DST.write_text(json.dumps(cfg, indent=1, sort_keys=True) + "\n")
    print(DST)
