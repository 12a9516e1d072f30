#!/usr/bin/env python3
"""Writes tests/golden/kats.json: known-answer vectors for the suffix-array
path, computed by pure-Python brute force (sorting suffix slices, the oracle
of the reference's own tests, sample_sort.rs:643-647 / 895-906 / 1021-1040).

Texts are the fixtures of the reference's tests: banana, mississippi,
abracadabra (ext_mem.rs:2784-2791), the unique-terminator text
(sample_sort.rs:814-821), the segmented "hello|world|banana|mississippi"
(sample_sort.rs:938-946), STAR "AAAA" (sample_sort.rs:1060-1070), and the LCG
texts of pack.rs:353-358 / 433-481 (seeds 12345, 999, 4242).  The reference
itself cannot run here (Rust, no toolchain), so these vectors come from the
same brute-force definition its tests use.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def lcg_texts():
    def lcg(state):
        state = (state * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
        return state, state >> 33

    out = {}
    st = 12345
    t = []
    for _ in range(4096):
        st, v = lcg(st)
        t.append(v % 4)
    out["lcg12345_plain"] = t
    for seed, period, cut in ((999, 617, 9), (4242, 411, 7)):
        st = seed
        t = []
        for i in range(8192):
            if i % period < cut:
                t.append(5)
            else:
                st, v = lcg(st)
                t.append(v % 5)
        out[f"lcg{seed}_spacer"] = t
    return out


def spacer_ends(text):
    """pack.rs:565-584."""
    ends, in_run = [], False
    for i, b in enumerate(text):
        if b == 5:
            if in_run:
                ends.append(i)
                in_run = False
        else:
            in_run = True
    if not ends or ends[-1] != len(text):
        ends.append(len(text))
    return ends


def lim_fn(n, ends):
    import bisect

    if ends is None:
        return lambda p: n - p
    return lambda p: (ends[bisect.bisect_right(ends, p)] - p) if bisect.bisect_right(ends, p) < len(ends) else n - p


def brute(text, positions, ends=None, star=False):
    n = len(text)
    lim = lim_fn(n, ends)

    def key(p):
        l = lim(p)
        body = tuple(text[p:p + l])
        # shorter-first: a prefix sorts before its extensions; STAR: longer
        # limit first, then position.  Ties (Equal) break by position.
        return body, l, p

    import functools

    def cmp(a, b):
        ka, kb = key(a), key(b)
        la, lb = ka[1], kb[1]
        m = min(la, lb)
        for i in range(m):
            if ka[0][i] != kb[0][i]:
                return -1 if ka[0][i] < kb[0][i] else 1
        if star:
            if la != lb:
                return -1 if la > lb else 1
            return (a > b) - (a < b)
        if la != lb:
            return -1 if la < lb else 1
        return (a > b) - (a < b)

    return sorted(positions, key=functools.cmp_to_key(cmp))


def main():
    cases = []
    for name, t in [("banana", b"banana"), ("mississippi", b"mississippi"), ("abracadabra", b"abracadabra"),
                    ("single", bytes([7])), ("empty", b""), ("unique_term", bytes([0, 1, 2, 0, 1, 5, 0, 2, 1, 6])),
                    ("run36", b"a" * 36), ("ab14", b"ab" * 14)]:
        text = list(t)
        cases.append({"name": name, "text": text, "ends": None, "order": 0, "positions": None,
                      "sa": brute(text, range(len(text)))})
    hw = list(b"helloworldbananamississippi")
    hw_ends = [5, 10, 16, 27]
    cases.append({"name": "segmented_hw", "text": hw, "ends": hw_ends, "order": 0, "positions": None,
                  "sa": brute(hw, range(len(hw)), hw_ends)})
    cases.append({"name": "star_hw", "text": hw, "ends": hw_ends, "order": 1, "positions": None,
                  "sa": brute(hw, range(len(hw)), hw_ends, star=True)})
    cases.append({"name": "star_AAAA", "text": list(b"AAAA"), "ends": [4], "order": 1, "positions": None,
                  "sa": [0, 1, 2, 3]})
    ms = list(b"mississippi")
    cases.append({"name": "mississippi_even", "text": ms, "ends": None, "order": 0, "positions": list(range(0, 11, 2)),
                  "sa": brute(ms, range(0, 11, 2))})
    for name, t in lcg_texts().items():
        if name.endswith("plain"):
            cases.append({"name": name, "text": t, "ends": None, "order": 0, "positions": None,
                          "sa": brute(t, range(len(t)))})
        else:
            ends = spacer_ends(t)
            acgt = [p for p in range(len(t)) if t[p] < 4]
            cases.append({"name": name + "_shorter", "text": t, "ends": ends, "order": 0, "positions": acgt,
                          "sa": brute(t, acgt, ends)})
            cases.append({"name": name + "_star", "text": t, "ends": ends, "order": 1, "positions": acgt,
                          "sa": brute(t, acgt, ends, star=True)})
    kats = {
        # lcp.rs:438-471
        "lcp": [["banana", 0, 1, None, 0], ["banana", 1, 3, None, 3], ["aaaaaa", 0, 1, 3, 3], ["abc", 0, 2, None, 0],
                ["abc", 1, 1, None, 2]],
        # lcp.rs:458-471 (-1 Less, 0 Equal, 1 Greater)
        "cmp": [["banana", 1, 0, -1], ["banana", 3, 1, -1], ["banana", 1, 1, 0]],
        # limits.rs:360-380: ends [3, 8, 15]
        "limits": {"ends": [3, 8, 15], "n": 15,
                   "lim": [[0, 3], [1, 2], [2, 1], [3, 5], [5, 3], [7, 1], [8, 7], [14, 1], [15, 0]]},
    }
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump({"cases": cases, "kats": kats}, f)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
