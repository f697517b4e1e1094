"""Regenerate the 10x10 matrices of reference_vectors.json from the reference's test sources.

The reference keeps its known-answer vectors in GoogleTest files (KAIZN-io/2DTissue,
tests/simulation/). This script reads test_Locomotion.cpp as text, pulls the numbers of the
`expected_diff_x` / `expected_diff_y` matrices of GetDistVectTestTenDimensionTest (:229-258) and
rewrites the "ten" entry of reference_vectors.json. Run it by hand where the reference is
mounted; tests only read the JSON.

    python tests/golden/extract_reference_vectors.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def numbers_after(text: str, marker: str, count: int) -> list[float]:
    start = text.index(marker) + len(marker)
    body = text[start:]
    body = body[: body.index(";")]
    body = re.sub(r"//", " ", body)
    vals = [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?(?:e-?\d+)?", body)]
    if len(vals) != count:
        raise ValueError(f"{marker}: expected {count} numbers, found {len(vals)}")
    return vals


def main(ref_root: str) -> None:
    src = open(os.path.join(ref_root, "tests/simulation/test_Locomotion.cpp")).read()
    ten = src[src.index("GetDistVectTestTenDimensionTest"):]
    dx = numbers_after(ten, "expected_diff_x <<", 100)
    dy = numbers_after(ten, "expected_diff_y <<", 100)
    path = os.path.join(HERE, "reference_vectors.json")
    d = json.load(open(path))
    t = d["get_dist_vect"]["ten"]
    t["dx"] = [dx[10 * i:10 * i + 10] for i in range(10)]
    t["dy"] = [dy[10 * i:10 * i + 10] for i in range(10)]
    for k in ("dx_row0", "dy_row0", "dx_row9", "dy_row9"):
        t.pop(k, None)
    json.dump(d, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
