"""Copy a covering subset of the reference's golden Mars 1 `.ifs` fixtures
(fixtures/mars1/*.ifs, data files the reference's own tests read) and their
manifest entries into tests/golden/mars1/.

Run in a checkout that has the reference at /root/reference.  Keeps, for each
test image, its fisher r8 default fixture, plus every non-default parameter
variant (alfa5, beta6, max32, min2, step8, the walkthrough) and the other two
domain-classification methods on three images.  The manifest values kept per
fixture are the ones the 1998 encoder and decoder printed: transform count,
zero-alfa transform count and the SHA-256 of `decmars -i`'s PGM output
(10 iterations, no postprocessing).
"""

import json
import re
import shutil
from pathlib import Path

SRC = Path("/root/reference/fixtures/mars1")
DST = Path(__file__).resolve().parent / "mars1"


def main():
    man = (SRC / "manifest.toml").read_text()
    keep = {}
    for block in man.split("[[fixture]]")[1:]:
        get = lambda k: re.search(rf'\n{k} = "?([^"\n]+)"?', "\n" + block).group(1)
        stem, image, variant, method = get("stem"), get("image"), get("variant"), get("method")
        t_rms = get("t_rms")
        pick = ((variant == "default" and method == "fisher" and t_rms == "8")
                or variant != "default"
                or (image in ("lena", "mixed_129x127", "zoneplate") and t_rms == "8"))
        if not pick:
            continue
        keep[stem] = {"file": get("file"), "width": int(get("width")), "height": int(get("height")),
                      "transforms": int(get("transforms")),
                      "zero_alfa_transforms": int(get("zero_alfa_transforms")),
                      "decode_iterative_sha256": get("decode_iterative_sha256"),
                      "ifs_sha256": get("ifs_sha256")}
    DST.mkdir(parents=True, exist_ok=True)
    for stem, rec in keep.items():
        shutil.copyfile(SRC / rec["file"], DST / rec["file"])
    (DST / "manifest.json").write_text(json.dumps(keep, indent=1, sort_keys=True) + "\n")
    print(f"{len(keep)} fixtures")


if __name__ == "__main__":
    main()
