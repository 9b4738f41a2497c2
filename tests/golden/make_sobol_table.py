"""Writes tests/golden/sobol_directions.json: the 128 Sobol direction words the
reference holds in shaders/sobol_table.glsl (dimensions 0..3, bit b of the
index at [d * 32 + b]), read from that file's text as data.

    python tests/golden/make_sobol_table.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (ref / "shaders" / "sobol_table.glsl").read_text()
    body = text[text.index("uint[128](") + len("uint[128]("):]
    body = body[:body.index(");")]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]{8})u", body)]
    assert len(words) == 128, len(words)
    out = {"source": "shaders/sobol_table.glsl:3-8", "layout": "word d * 32 + b: dimension d, index bit b",
           "directions": [f"0x{w:08x}" for w in words]}
    (HERE / "sobol_directions.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(words)} direction words")


if __name__ == "__main__":
    main()
