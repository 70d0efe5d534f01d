"""Writes sampler_states.json: the first 8 sampler states of a few
(pass, pixelId) pairs, from a pure-Python restatement of the reference sampler
(sampling/random_sampler.h:44-73, math/hash.h:23-72).  Not produced by the
oracle it checks."""
import json
import os

M = 0xFFFFFFFF


def rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & M


def mix(h, k):
    k = (k * 0xCC9E2D51) & M
    k = rotl(k, 15)
    k = (k * 0x1B873593) & M
    h ^= k
    return (rotl(h, 13) * 5 + 0xE6546B64) & M


def fmix(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M
    return h ^ (h >> 16)


def states(p, pixel, n=8):
    s = fmix(mix(mix(0, pixel & M), p & M))
    out = []
    for _ in range(n):
        s = (s * 1664525 + 1013904223) & M
        out.append(s)
    return out


if __name__ == "__main__":
    pairs = [(0, 0), (0, 1), (1, 0), (5, 262143), (15, 2073599), (63, 8294399), (1000, 123456)]
    rows = [{"pass": p, "pixel": px, "states": states(p, px)} for p, px in pairs]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampler_states.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_sampler_golden.py", "rows": rows}, f, indent=1)
