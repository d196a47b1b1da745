"""Writes the small data fixtures under tests/golden/ from the reference's own
data files (run in a container where the reference checkout is present):

* precomputed_coloring.json — the reference's cached greedy edge colouring of a
  208-edge mesh (precomputed/concat_prefix_sum.npy, concat_sorted_edges_index.npy,
  read with numpy.load(allow_pickle=False));
* plane_8.obj — the 81-vertex / 128-face cloth asset (models/OBJ/plane_8.obj),
  the mesh whose 208 edges (V + F - 1) that colouring is of.
"""
import json
import os
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

pre = np.load(os.path.join(REF, "precomputed", "concat_prefix_sum.npy"), allow_pickle=False)
idx = np.load(os.path.join(REF, "precomputed", "concat_sorted_edges_index.npy"), allow_pickle=False)
with open(os.path.join(HERE, "precomputed_coloring.json"), "w") as fh:
    json.dump({"source": "precomputed/concat_prefix_sum.npy, precomputed/concat_sorted_edges_index.npy",
               "color_start": pre.astype(int).tolist(), "sorted_edges": idx.astype(int).tolist()}, fh)
shutil.copy(os.path.join(REF, "models", "OBJ", "plane_8.obj"), os.path.join(HERE, "plane_8.obj"))
print("fixtures written")
