"""Loading the transcribed hash-map data sets (tests/golden/hashmap_cases.json) as arrays."""
import json
import os

import numpy as np

import oracle

PATH = os.path.join(os.path.dirname(__file__), "golden", "hashmap_cases.json")


def load():
    with open(PATH) as f:
        return json.load(f)


def hashmap(spec):
    hm = np.zeros(spec["len"], oracle.HASH_ENTRY)
    for k, e in spec["entries"].items():
        hm[int(k)] = (e["current"], e["next"], e["steps"], e["predecessors"], e["state"])
    return hm


def u32_array(spec):
    a = np.zeros(spec["len"], np.uint32)
    for k, v in spec["values"].items():
        a[int(k)] = v
    return a
