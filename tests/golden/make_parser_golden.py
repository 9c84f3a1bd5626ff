#!/usr/bin/env python3
"""Write parser fixtures: the reference's example control files (examples/dmpci.* and the
dmpci files inside examples/Biology-2023-Input-Files.zip), each with the polymer, bead,
bond and bond-pair counts the compiled reference creates for it (its dmpcis summary,
src/InitialState.cpp:451-521).

Run in the development container only:
    make -f oracle/ref.mk -j8 && python tests/golden/make_parser_golden.py

To get the counts cheaply each input is run for one step: Time and the output periods
are set to 1 and the Process/Analysis/Event/Command blocks are dropped.  The fixture
keeps the original text, which the parser must read whole.
"""
import concurrent.futures as cf
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "dpd")
SRC = "/root/reference/examples"
OUT = os.path.join(HERE, "parser")

PERIODS = ("Time", "SamplePeriod", "AnalysisPeriod", "DensityPeriod", "DisplayPeriod", "RestartPeriod")


def one_step(text):
    lines = []
    for line in text.splitlines():
        tok = line.split()
        if tok and tok[0] in PERIODS:
            line = f"{tok[0]}\t1"
        lines.append(line)
    txt = "\n".join(lines) + "\n"
    # drop everything from the first trailing block keyword on
    m = re.search(r"^\s*(Process|Analysis|Event|Command)\b", txt[txt.index("Grid"):], re.M)
    if m:
        txt = txt[: txt.index("Grid") + m.start()]
    return txt


def counts(name, text):
    work = tempfile.mkdtemp(prefix="dpd_parse_")
    try:
        with open(os.path.join(work, "dmpci.p"), "w") as fh:
            fh.write(one_step(text))
        r = subprocess.run([REF, "p"], cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=900)
        path = os.path.join(work, "dmpcis.p")
        if not os.path.exists(path):
            return name, None
        txt = open(path).read()
        out = {}
        for key, pat in (("polymers", "Number of polymers created"), ("beads", "Number of beads created"),
                         ("bonds", "Number of bonds created"), ("bondpairs", "Number of bond pairs created")):
            m = re.search(pat + r"\s+(\d+)", txt)
            if not m:
                return name, None
            out[key] = int(m.group(1))
        return name, out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    os.makedirs(OUT, exist_ok=True)
    inputs = {}
    for f in sorted(os.listdir(SRC)):
        if f.startswith("dmpci."):
            inputs["examples_" + f[6:]] = open(os.path.join(SRC, f)).read()
    with zipfile.ZipFile(os.path.join(SRC, "Biology-2023-Input-Files.zip")) as z:
        for info in sorted(z.infolist(), key=lambda i: i.filename):
            base = os.path.basename(info.filename)
            if base.startswith("dmpci.") and not info.is_dir():
                key = "bio_" + os.path.dirname(info.filename).replace("/", "_") + "_" + base[6:]
                inputs[key] = z.read(info).decode()
    result = {}
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        for name, c in ex.map(lambda kv: counts(*kv), inputs.items()):
            result[name] = c
            print(name, c, flush=True)
    for name, text in inputs.items():
        with open(os.path.join(OUT, name + ".dmpci"), "w") as fh:
            fh.write(text)
    with open(os.path.join(OUT, "counts.json"), "w") as fh:
        json.dump(result, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
