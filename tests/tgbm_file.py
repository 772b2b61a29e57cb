"""TGBM v3 model files read into plain Python values and written back (test infrastructure).

A third statement of the file layout, independent of both C++ libraries, used to build
corrupted or edited files for the loader tests.  Field order follows the reference writer
(reference cpp/src/serialize.cpp:65-148).  `read` then `write` must give back the same bytes;
tests/test_model_file.py checks that on files the reference core wrote.
"""
from __future__ import annotations

import struct

CONFIG = [("n_estimators", "i"), ("learning_rate", "d"), ("max_depth", "i"), ("max_leaves", "i"),
          ("min_samples_leaf", "i"), ("min_child_weight", "d"), ("l2", "d"), ("min_split_gain", "d"),
          ("max_bin", "i"), ("subsample", "d"), ("colsample", "d"), ("n_time_buckets_cfg", "i"),
          ("horizon", "q"), ("embargo", "q"), ("iid_mode", "i"), ("halo", "i"), ("objective", "i"),
          ("quantile", "d"), ("huber_delta", "d"), ("rank_temperature", "d"), ("random_state", "i"),
          ("n_jobs", "i"), ("tree_decay", "d"), ("n_features", "i"), ("n_bins", "i"), ("n_time_buckets", "i"),
          ("best_iteration", "i"), ("init_score", "d")]
NODE = [("is_leaf", "i"), ("split_type", "i"), ("feature", "i"), ("feature2", "i"), ("threshold", "i"),
        ("left", "i"), ("right", "i"), ("missing_go_left", "i"), ("value", "d"), ("gain", "d"),
        ("n_samples", "i"), ("depth", "i"), ("cat_left", "4Q")]


class _In:
    def __init__(self, b: bytes):
        self.b, self.at = b, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.at)
        self.at += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def vec(self, fmt):
        n = self.take("i")
        return [self.take(fmt) for _ in range(n)]


def read(blob: bytes) -> dict:
    r = _In(blob)
    assert blob[:4] == b"TGBM"
    r.at = 4
    m = {"version": r.take("I")}
    for k, f in CONFIG:
        m[k] = r.take(f)
    m["init_prefix"] = r.vec("d")
    names = []
    for _ in range(r.take("i")):
        n = r.take("i")
        names.append(r.b[r.at:r.at + n].decode())
        r.at += n
    m["feature_names"] = names
    m["is_cat"] = r.vec("i")
    m["group_of"] = r.vec("i")
    m["monotone"] = r.vec("i")
    m["binner_max_bin"] = r.take("i")
    feats = []
    for _ in range(r.take("i")):
        feats.append({"is_cat": r.take("i"), "edges": r.vec("d"), "cat_values": r.vec("d"), "n_bins": r.take("i")})
    m["features"] = feats
    trees = []
    for _ in range(r.take("i")):
        t = {"n_buckets": r.take("i"), "nodes": []}
        for _ in range(r.take("i")):
            node = {}
            for k, f in NODE:
                v = r.take(f)
                node[k] = list(v) if f == "4Q" else v
            t["nodes"].append(node)
        t["values"] = r.vec("d")
        t["cover"] = r.vec("i")
        t["prefix_values"] = r.vec("d")
        t["target"] = r.take("i")
        t["n_targets"] = r.take("i")
        trees.append(t)
    m["trees"] = trees
    m["objective_kind"] = r.take("i")
    assert r.at == len(blob), "trailing bytes"
    return m


def write(m: dict) -> bytes:
    out = [b"TGBM", struct.pack("<I", m["version"])]
    put = lambda f, v: out.append(struct.pack("<" + f, *(v if isinstance(v, (list, tuple)) else [v])))

    def vec(f, xs):
        put("i", len(xs))
        for x in xs:
            put(f, x)

    for k, f in CONFIG:
        put(f, m[k])
    vec("d", m["init_prefix"])
    put("i", len(m["feature_names"]))
    for s in m["feature_names"]:
        b = s.encode()
        put("i", len(b))
        out.append(b)
    vec("i", m["is_cat"])
    vec("i", m["group_of"])
    vec("i", m["monotone"])
    put("i", m["binner_max_bin"])
    put("i", len(m["features"]))
    for ft in m["features"]:
        put("i", ft["is_cat"])
        vec("d", ft["edges"])
        vec("d", ft["cat_values"])
        put("i", ft["n_bins"])
    put("i", len(m["trees"]))
    for t in m["trees"]:
        put("i", t["n_buckets"])
        put("i", len(t["nodes"]))
        for node in t["nodes"]:
            for k, f in NODE:
                put(f, node[k])
        vec("d", t["values"])
        vec("i", t["cover"])
        vec("d", t["prefix_values"])
        put("i", t["target"])
        put("i", t["n_targets"])
    put("i", m["objective_kind"])
    return b"".join(out)
