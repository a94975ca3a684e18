"""Decode the reference's HDF5 test databases into plain .npz fixtures.

Run in a container that has h5dump (HDF5 command-line tools); the outputs are
committed so tests can check the library's own HDF5 reader against an
independent decoding without needing h5dump.

  small_mass.npz        names, kmers, sketchsize64, bbits, sketches [n][K][s64*bbits]
  small_mass_plus1.npz  the same for the 7-sample database

The .h5 files are the reference's own test fixtures (test/small_mass*.h5),
copied as data.
"""
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")
H5LS = os.environ.get("H5LS", "/opt/conda/bin/h5ls")


def _run(args):
    return subprocess.run(args, check=True, stdout=subprocess.PIPE, text=True).stdout


def _values(text):
    body = text[text.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    body = re.sub(r"\(\d+\):", " ", body)
    return [int(x) for x in re.findall(r"-?\d+", body)]


def decode(path):
    names = [l.split()[0] for l in _run([H5LS, path + "/sketches"]).splitlines() if l.strip()]
    names.sort()
    kmers = None
    s64 = bbits = None
    sk = []
    for nm in names:
        grp = "/sketches/" + nm
        k = _values(_run([H5DUMP, "-a", grp + "/kmers", path]))
        s = _values(_run([H5DUMP, "-a", grp + "/sketchsize64", path]))[0]
        b = _values(_run([H5DUMP, "-a", grp + "/bbits", path]))[0]
        if kmers is None:
            kmers, s64, bbits = k, s, b
        assert k == kmers and s == s64 and b == bbits
        per_k = []
        for kk in kmers:
            v = _values(_run([H5DUMP, "-w", "0", "-d", grp + "/" + str(kk), path]))
            assert len(v) == s64 * bbits
            per_k.append(np.array(v, dtype=np.uint64))
        sk.append(np.stack(per_k))
    return (np.array(names), np.array(kmers, np.int32), s64, bbits,
            np.stack(sk).astype(np.uint64))


def main():
    for base in ("small_mass", "small_mass_plus1"):
        names, kmers, s64, bbits, sk = decode(os.path.join(HERE, base + ".h5"))
        np.savez_compressed(os.path.join(HERE, base + ".npz"), names=names, kmers=kmers,
                            sketchsize64=np.int64(s64), bbits=np.int64(bbits), sketches=sk)
        print(base, names.tolist(), kmers.tolist(), sk.shape, file=sys.stderr)


if __name__ == "__main__":
    main()
