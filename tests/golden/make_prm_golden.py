"""Writes tests/golden/prm_parser.json: what the reference's own .prm parser
(Src/FileParams.cpp, built from the reference sources into
oracle/_ref/ref_fileparams by oracle/ref.mk) prints for a set of .prm files,
and its exit status.  The product parser (`LISTCTRecons --prm`) is checked
against these vectors wherever the reference sources are absent.

    make -C oracle -f ref.mk && python tests/golden/make_prm_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from listctrecons_amd import scan  # noqa: E402

REF_PARSER = os.path.join(ROOT, "oracle", "_ref", "ref_fileparams")


def cases():
    out = {}
    for name, w in (
            ("tiny_offsets", scan.workload("tiny", view_offset=3, data_offset=5)),
            ("plain_256x32", scan.workload("helical_16x368_256x32")),
            ("clinical_512x256", scan.workload("helical_64x736_512x256")),
            ("large_1024x512", scan.workload("helical_96x896_1024x512", offset_x=1.25,
                                              offset_y=-0.5)),
            ("pitch15_256x128", scan.workload("helical_16x368_p15_256x128")),
    ):
        out[name] = w.prm_text("/data/scans/" + name + "/", "/data/out/")
    base = out["tiny_offsets"]
    # the alias key, extra unknown keys, one line, tabs and CRLF line ends
    out["tablefeed_alias"] = base.replace("PitchValue:", "TableFeed:")
    out["extra_tokens"] = "Comment: hello\n" + base + "SmoothWindow: hann\nReconType: x\n"
    out["one_line"] = " ".join(base.split())
    out["tabs"] = base.replace(": ", ":\t")
    out["crlf"] = base.replace("\n", "\r\n")
    # a required key missing (the reference exits 1)
    out["missing_dz"] = "".join(l + "\n" for l in base.splitlines() if not l.startswith("Dz:"))
    out["missing_nprojturn"] = "".join(l + "\n" for l in base.splitlines()
                                       if not l.startswith("NProjTurn:"))
    return out


def main():
    if not os.path.exists(REF_PARSER):
        sys.exit("build the reference parser first: make -C oracle -f ref.mk")
    vectors = {}
    with tempfile.TemporaryDirectory() as d:
        for name, text in cases().items():
            p = os.path.join(d, name + ".prm")
            with open(p, "w", newline="") as f:
                f.write(text)
            r = subprocess.run([REF_PARSER, p], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True)
            vectors[name] = {"prm": text, "stdout": r.stdout, "returncode": r.returncode}
    with open(os.path.join(HERE, "prm_parser.json"), "w") as f:
        json.dump(vectors, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {len(vectors)} vectors")


if __name__ == "__main__":
    main()
