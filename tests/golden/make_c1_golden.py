"""Writes tests/golden/c1_seed0.json: SHA-256 of the 1 MiB seed-0 document and of
the oracle's outputs on it (line offsets, nodes, tree links).  A regression pin
for the generator and the oracle; the parity tests compare the HIP path against
the oracle on this same document."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oneline_amd import datagen  # noqa: E402
from oracle import pyoracle as orc  # noqa: E402

text = datagen.document(1 << 20, seed=0)
off, nodes = orc.parse(text)
links = orc.links(nodes)
g = {
    "workload": "1 MiB synthetic document, seed 0",
    "text_sha256": hashlib.sha256(text.tobytes()).hexdigest(),
    "lines": int(len(off)),
    "offsets_sha256": hashlib.sha256(off.tobytes()).hexdigest(),
    "nodes_sha256": hashlib.sha256(nodes.tobytes()).hexdigest(),
}
for name, arr in zip(("parent", "subtree_end", "first_child", "next_sibling"), links):
    g[name + "_sha256"] = hashlib.sha256(arr.tobytes()).hexdigest()
with open(os.path.join(ROOT, "tests", "golden", "c1_seed0.json"), "w") as f:
    json.dump(g, f, indent=1)
print(json.dumps(g, indent=1))
