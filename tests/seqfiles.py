"""Sequence files for the sketching tests: genomes evolved from one ancestor
(SNPs at a per-genome rate plus an accessory block each), written as plain,
gzipped, wrapped, CRLF, lower-case and N-holding FASTA, split over two
files, and as FASTQ reads."""
import gzip

import numpy as np


def genomes(rng, n, length):
    anc = rng.choice(list("ACGT"), size=length)
    out = []
    for i in range(n):
        g = anc.copy()
        m = rng.random(length) < 0.004 * (i + 1)
        g[m] = rng.choice(list("ACGT"), size=int(m.sum()))
        acc = rng.choice(list("ACGT"), size=int(rng.integers(2000, 8000)))
        out.append("".join(g) + "".join(acc))
    return out


def fasta(records, wrap=0, crlf=False):
    nl = "\r\n" if crlf else "\n"
    parts = []
    for name, s in records:
        parts.append(f">{name} desc{nl}")
        if wrap:
            parts.extend(s[i:i + wrap] + nl for i in range(0, len(s), wrap))
        else:
            parts.append(s + nl)
    return "".join(parts)


def put(path, text, gz=False):
    data = text.encode()
    with (gzip.open(path, "wb") if gz else open(path, "wb")) as f:
        f.write(data)
    return str(path)


def samples(tmp_path, seed=4):
    rng = np.random.default_rng(seed)
    g = genomes(rng, 6, 60000)
    files = []
    files.append([put(tmp_path / "s0.fa", fasta([("c1", g[0])]))])
    mid = len(g[1]) // 2
    files.append([put(tmp_path / "s1.fa.gz",
                       fasta([("c1", g[1][:mid]), ("tiny", g[1][mid:mid + 20]),
                               ("c2", g[1][mid:].lower())]), gz=True)])
    s2 = list(g[2])
    for p in rng.integers(0, len(s2) - 50, size=10):
        s2[p:p + 30] = "N" * 30
    files.append([put(tmp_path / "s2.fa", fasta([("c1", "".join(s2))], wrap=60, crlf=True))])
    files.append([put(tmp_path / "s3a.fa", fasta([("c1", g[3][:30000])])),
                  put(tmp_path / "s3b.fa.gz", fasta([("c2", g[3][30000:])]), gz=True)])
    reads = []
    for r in range(2000):
        p = int(rng.integers(0, len(g[4]) - 150))
        reads.append(f"@r{r}\n{g[4][p:p + 150]}\n+\n{'I' * 150}\n")
    files.append([put(tmp_path / "s4.fq.gz", "".join(reads), gz=True)])
    files.append([put(tmp_path / "s5.fa", fasta([("c1", g[5])], wrap=80))])
    names = [f"sample_{i}" for i in range(len(files))]
    return names, files


def fastq_reads(rng, genome, n_reads, read_len=100, error=0.01, n_rate=0.0):
    """Reads sampled uniformly from genome (str), either strand, with
    substitution errors and optional N calls; FASTQ text."""
    comp = str.maketrans("ACGT", "TGCA")
    out = []
    for r in range(n_reads):
        p = int(rng.integers(0, len(genome) - read_len + 1))
        s = genome[p:p + read_len]
        if rng.random() < 0.5:
            s = s.translate(comp)[::-1]
        s = np.array(list(s))
        m = rng.random(read_len) < error
        s[m] = rng.choice(list("ACGT"), size=int(m.sum()))
        if n_rate:
            s[rng.random(read_len) < n_rate] = "N"
        out.append(f"@r{r}\n{''.join(s)}\n+\n{'I' * read_len}\n")
    return "".join(out)
