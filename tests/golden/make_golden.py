"""Writes the golden fixtures of tests/golden/ (run from the repository root).

sort_cases.json
    The literal sort cases of the reference's own tests (crates/dynamis-sort/tests/gpu.rs):
    inputs, word counts, and the expected outputs. Expected orders come from Python's stable
    `sorted` on the masked key bytes (a textbook stable sort, independent of the oracle), and
    where the reference test spells its expectation out literally, that literal is kept too.
    Also: the key_words table of gpu.rs:313-320 and the first outputs of the xorshift generator
    of gpu.rs:214-224.
grid_cases.json
    Known-answer cell hashes and keys worked by hand from grid_index.wgsl:28-37 (the derivation
    of each value is spelled out next to it), and the broadphase scenes of
    crates/dynamis-world/tests/suite/broadphase.rs:297-352 restated as swept boxes, with the pair
    counts those tests assert.
"""

from __future__ import annotations

import json
import pathlib

HERE = pathlib.Path(__file__).resolve().parent


def xorshift_keys(n: int, salt: int) -> list[int]:
    """gpu.rs:214-224."""
    state = (0x9E3779B9 ^ salt) & 0xFFFFFFFF
    out = []
    for _ in range(n):
        state ^= (state << 13) & 0xFFFFFFFF
        state ^= state >> 17
        state ^= (state << 5) & 0xFFFFFFFF
        out.append(state)
    return out


def stable_expected(major, minor, payload, major_words, minor_words):
    mmask = (1 << (8 * major_words)) - 1 if major_words else 0
    nmask = (1 << (8 * minor_words)) - 1 if minor_words else 0
    idx = sorted(range(len(major)), key=lambda i: (major[i] & mmask, minor[i] & nmask))
    return ([major[i] for i in idx], [minor[i] for i in idx], [payload[i] for i in idx])


def sort_cases() -> dict:
    cases = []

    def add(name, source, major, minor, payload, major_words, minor_words, literal=None):
        em, en, ep = stable_expected(major, minor, payload, major_words, minor_words)
        case = {"name": name, "source": source, "major": major, "minor": minor,
                "payload": payload, "major_words": major_words, "minor_words": minor_words,
                "expected_major": em, "expected_minor": en, "expected_payload": ep}
        if literal:
            case["literal"] = literal
        cases.append(case)

    keys = [3, 5, 1, 0, 7, 2, 2, 9, 4, 6]
    add("stable_with_payload", "crates/dynamis-sort/tests/gpu.rs:192-212", keys, keys,
        list(range(10)), 1, 0)
    n = 64
    add("major_above_minor", "crates/dynamis-sort/tests/gpu.rs:252-274",
        [i % 4 for i in range(n)], [(i * 5) % 8 for i in range(n)], list(range(n)), 1, 1)
    keys = [1, 1, 0, 0, 2, 2]
    mx = 0xFFFFFFFF
    add("payload_never_keys", "crates/dynamis-sort/tests/gpu.rs:277-288", keys, keys,
        [mx, 0, mx, 0, mx, 0], 1, 0,
        literal={"major": [0, 0, 1, 1, 2, 2], "payload": [mx, 0, mx, 0, mx, 0]})
    keys = [2, 0, 1, 2, 0, 1, 3, 3, 0, 2]
    add("equal_keys_input_order", "crates/dynamis-sort/tests/gpu.rs:291-310", keys, keys,
        list(range(10)), 1, 0)
    keys = [9, 8, 7, 6, 5, 4, 3, 2, 1, 0]
    add("rebound_lane", "crates/dynamis-sort/tests/gpu.rs:369-380", keys, keys,
        list(range(10)), 1, 0)
    return {
        "cases": cases,
        "key_words": [[1, 1], [256, 1], [257, 2], [1 << 16, 2], [1 << 24, 3],
                      [0xFFFFFFFF, 4]],
        "key_words_source": "crates/dynamis-sort/tests/gpu.rs:313-320",
        "xorshift_salt1_first8": xorshift_keys(8, 1),
        "xorshift_source": "crates/dynamis-sort/tests/gpu.rs:214-224",
    }


def grid_cases() -> dict:
    # Hand derivations (32-bit wrapping arithmetic, then & 0x07FFFFFF):
    #   (0,0,0): every product is 0                                   -> 0x00000000
    #   (1,0,0): x = 0x9E3779B9, x<<7 = 0x1BBCDC80 (low 32 bits);
    #            0x9E3779B9 ^ 0x1BBCDC80 = 0x858BA539                  -> 0x058BA539
    #   (0,1,0): y = 0x85EBCA77, y>>3 = 0x10BD794E;
    #            0x85EBCA77 ^ 0x10BD794E = 0x9556B339                  -> 0x0556B339
    #   (0,0,1): z = 0xC2B2AE3D, z<<11 = 0x95571E800 -> 0x9571E800;
    #            0xC2B2AE3D ^ 0x9571E800 = 0x57C3463D                  -> 0x07C3463D
    #   (-1,0,0): x = 0xFFFFFFFF * 0x9E3779B9 = -0x9E3779B9 = 0x61C88647,
    #            x<<7 = 0xE4432380; 0x61C88647 ^ 0xE4432380 = 0x858BA5C7 -> 0x058BA5C7
    hashes = [
        {"coord": [0, 0, 0], "hash": 0x00000000},
        {"coord": [1, 0, 0], "hash": 0x058BA539},
        {"coord": [0, 1, 0], "hash": 0x0556B339},
        {"coord": [0, 0, 1], "hash": 0x07C3463D},
        {"coord": [-1, 0, 0], "hash": 0x058BA5C7},
    ]
    # cell_key(level, c) = (level << 27) | hash(c)   (grid_index.wgsl:35-37)
    keys = [
        {"level": 3, "coord": [1, 0, 0], "key": (3 << 27) | 0x058BA539},
        {"level": 30, "coord": [0, 0, 0], "key": 30 << 27},
    ]
    margin = 0.02  # PhysicsConfig::default contact_margin (config.rs:76-100)

    def sphere(center, r, moving):
        grow = r + (margin if moving else 0.0)
        tight = ([c - r for c in center], [c + r for c in center])
        swept = ([c - grow for c in center], [c + grow for c in center])
        return {"min": swept[0], "max": swept[1], "res_min": tight[0], "res_max": tight[1],
                "mobile": moving}

    def cuboid(center, half):
        lo = [c - h for c, h in zip(center, half)]
        hi = [c + h for c, h in zip(center, half)]
        return {"min": lo, "max": hi, "res_min": lo, "res_max": hi, "mobile": False}

    scenes = []
    for offset in [0.05, 0.2, 0.35, 0.5, 0.65]:
        scenes.append({
            "name": f"two_spheres_offset_{offset}",
            "source": "crates/dynamis-world/tests/suite/broadphase.rs:297-313",
            "bodies": [sphere([0.0, 0.0, 0.0], 0.35, False), sphere([offset, 0.0, 0.0], 0.35, True)],
            "pairs": 1,
        })
    scenes.append({
        "name": "chain_of_16",
        "source": "crates/dynamis-world/tests/suite/broadphase.rs:315-329",
        "bodies": [sphere([i * 0.5, 0.0, 0.0], 0.4, True) for i in range(16)],
        "pairs": 15,
    })
    crowd = [cuboid([0.0, -0.5, 0.0], [30.0, 0.5, 30.0])]
    for index in range(64):
        column = (index % 8) - 3.5
        row = (index // 8) - 3.5
        crowd.append(sphere([column, 0.4, row], 0.4, True))
    scenes.append({
        "name": "floor_and_64_spheres",
        "source": "crates/dynamis-world/tests/suite/broadphase.rs:331-352",
        "bodies": crowd,
        "pairs": 64,
    })
    return {"hashes": hashes, "keys": keys, "scenes": scenes,
            "source": "crates/dynamis-shader/shaders/grid_index.wgsl:28-37"}


def main() -> None:
    (HERE / "sort_cases.json").write_text(json.dumps(sort_cases(), indent=1) + "\n")
    (HERE / "grid_cases.json").write_text(json.dumps(grid_cases(), indent=1) + "\n")


if __name__ == "__main__":
    main()
