"""Determinant and wavefunction files through the drop-in C++ API (include/sbd_amd/sbd.h):
LoadAlphaDets on .txt and binary .bin files (makedeterminants.h:63-105, :160-185) and
SaveMatrixFormWF (restart.h:662-780).  Host code only.  The expected words are computed here
from the strings themselves (a string read right to left is orbital 0 upwards; a .bin record
is that number in big-endian bytes); where the compiled reference driver is built
(oracle/_ref/sbd_ref), its files for the same inputs must be identical byte for byte."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
REF = os.path.join(ROOT, "oracle", "_ref", "sbd_ref")
BL = 20


@pytest.fixture(scope="module")
def shim_io():
    if not os.path.exists("/opt/conda/include/mpi.h") or shutil.which("g++") is None:
        pytest.skip("needs g++ and the image's MPI to build tests/cpp/shim_io")
    r = subprocess.run(["make", "-s", "-C", CPP, "shim_io"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return os.path.join(CPP, "shim_io")


def strings(norb, nel, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        occ = rng.choice(norb, nel, replace=False)
        out.append(sum(1 << int(p) for p in occ))
    out += out[: count // 5]  # duplicates, removed by the loader
    rng.shuffle(out)
    return out


def words(values, norb):
    nw = (norb + BL - 1) // BL
    return np.array([[(v >> (BL * w)) & ((1 << BL) - 1) for w in range(nw)] for v in sorted(set(values))],
                    dtype=np.uint64).ravel()


def write_txt(path, values, norb):
    with open(path, "w") as f:
        for v in values:
            f.write(format(v, "b").zfill(norb) + "\n")


def write_bin(path, values, norb, pad_bits=False, tail=False):
    nbyte = (norb + 7) // 8
    with open(path, "wb") as f:
        for v in values:
            rec = bytearray(v.to_bytes(nbyte, "big"))
            if pad_bits and 8 * nbyte > norb:
                rec[0] |= (0xFF << (8 - (8 * nbyte - norb))) & 0xFF  # padding above orbital norb - 1
            f.write(bytes(rec))
        if tail:
            f.write(b"\x01")  # a partial record, ignored


CASES = [(12, 4, 60), (36, 8, 300), (48, 12, 200), (64, 20, 150)]


@pytest.mark.parametrize("norb,nel,count", CASES)
def test_txt_and_bin_dets_load_the_same_sorted_words(shim_io, tmp_path, norb, nel, count):
    a = strings(norb, nel, count, seed=norb)
    b = strings(norb, nel, count, seed=norb + 1)
    write_txt(tmp_path / "a.txt", a, norb)
    write_bin(tmp_path / "b.BIN", b, norb, pad_bits=True, tail=True)
    out = str(tmp_path / "shim")
    r = subprocess.run([shim_io, str(tmp_path / "a.txt"), str(tmp_path / "b.BIN"), str(norb), str(BL), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    got_a = np.fromfile(out + ".adet.u64", dtype=np.uint64)
    got_b = np.fromfile(out + ".bdet.u64", dtype=np.uint64)
    assert np.array_equal(got_a, words(a, norb))
    assert np.array_equal(got_b, words(b, norb))
    # the same strings as .bin and as .txt load identically
    write_bin(tmp_path / "a.bin", a, norb)
    write_txt(tmp_path / "b.txt", b, norb)
    out2 = str(tmp_path / "shim2")
    r = subprocess.run([shim_io, str(tmp_path / "a.bin"), str(tmp_path / "b.txt"), str(norb), str(BL), out2],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    assert np.array_equal(np.fromfile(out2 + ".adet.u64", dtype=np.uint64), got_a)
    assert np.array_equal(np.fromfile(out2 + ".bdet.u64", dtype=np.uint64), got_b)


def test_matrix_form_wavefunction_files(shim_io, tmp_path):
    norb, nel = 10, 3
    a = strings(norb, nel, 12, seed=5)
    b = strings(norb, nel, 9, seed=6)
    write_txt(tmp_path / "a.txt", a, norb)
    write_bin(tmp_path / "b.bin", b, norb)
    na, nb = len(set(a)), len(set(b))
    W = np.random.default_rng(7).standard_normal(na * nb) * 10.0 ** np.arange(-8, 8, 16 / (na * nb))[: na * nb]
    W[3] = 0.0
    W.tofile(tmp_path / "w.f64")
    out = str(tmp_path / "shim")
    r = subprocess.run([shim_io, str(tmp_path / "a.txt"), str(tmp_path / "b.bin"), str(norb), str(BL), out,
                        str(tmp_path / "w.f64")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    assert open(out + ".wf.bin", "rb").read() == W.tobytes()
    lines = open(out + ".wf.txt").read().splitlines()
    assert lines == ["%g # %d %d" % (W[q], q // nb, q % nb) for q in range(na * nb)]


def test_unknown_extension_loads_nothing(shim_io, tmp_path):
    write_txt(tmp_path / "a.dets", [7, 11], 6)
    write_txt(tmp_path / "b.txt", [7, 11], 6)
    out = str(tmp_path / "shim")
    r = subprocess.run([shim_io, str(tmp_path / "a.dets"), str(tmp_path / "b.txt"), "6", str(BL), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    assert '"na":0' in r.stdout and '"nb":2' in r.stdout


def test_missing_bin_file_raises(shim_io, tmp_path):
    write_txt(tmp_path / "b.txt", [7, 11], 6)
    r = subprocess.run([shim_io, str(tmp_path / "none.bin"), str(tmp_path / "b.txt"), "6", str(BL),
                        str(tmp_path / "o")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Failed to open file" in r.stderr


@pytest.mark.skipif(not os.path.exists(REF), reason="compiled reference driver not built")
@pytest.mark.parametrize("norb,nel,count", CASES[:3])
def test_files_identical_to_the_compiled_reference(shim_io, tmp_path, norb, nel, count):
    a = strings(norb, nel, count, seed=norb + 10)
    b = strings(norb, nel, count // 2, seed=norb + 11)
    write_bin(tmp_path / "a.bin", a, norb, pad_bits=True, tail=True)
    write_txt(tmp_path / "b.txt", b, norb)
    W = np.random.default_rng(norb).standard_normal(len(set(a)) * len(set(b)))
    W.tofile(tmp_path / "w.f64")
    mine, ref = str(tmp_path / "shim"), str(tmp_path / "ref")
    r = subprocess.run([shim_io, str(tmp_path / "a.bin"), str(tmp_path / "b.txt"), str(norb), str(BL), mine,
                        str(tmp_path / "w.f64")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([REF, "--op", "files", "--adets", str(tmp_path / "a.bin"), "--bdets",
                        str(tmp_path / "b.txt"), "--norb", str(norb), "--bit_length", str(BL), "--out", ref,
                        "--vec", str(tmp_path / "w.f64")], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    for ext in (".adet.u64", ".bdet.u64", ".wf.bin", ".wf.txt"):
        assert open(mine + ext, "rb").read() == open(ref + ext, "rb").read(), ext


@pytest.mark.parametrize("norb,bl", [(6, 20), (12, 7), (36, 20), (40, 64)])
def test_load_from_alpha_dets_configurations(shim_io, tmp_path, norb, bl):
    """LoadFromAlphaDets (makedeterminants.h:110-149, one rank): the file's strings in file
    order (no sorting, duplicates kept) serve as both spins; configuration n * ib + ia holds
    string ia's norb bits then string ib's, repacked at bl bits per word.  Expected words are
    built here from the integers written to the file."""
    rng = np.random.default_rng(norb)
    vals = [int(v) for v in rng.integers(0, 1 << min(norb, 62), size=9)]
    vals.append(vals[2])  # a duplicate stays
    # a one-byte tail is a whole record when records are one byte long
    write_bin(tmp_path / "d.bin", vals, norb, pad_bits=True, tail=norb > 8)
    out = str(tmp_path / "cfg")
    r = subprocess.run([shim_io, "configs", str(tmp_path / "d.bin"), str(norb), str(bl), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    n = len(vals)
    nw = (2 * norb + bl - 1) // bl
    want = []
    for ib in range(n):
        for ia in range(n):
            full = vals[ia] | (vals[ib] << norb)
            want += [(full >> (bl * w)) & ((1 << bl) - 1) for w in range(nw)]
    got = np.fromfile(out + ".config.u64", dtype=np.uint64)
    assert got.size == n * n * nw
    assert np.array_equal(got, np.array(want, dtype=np.uint64))


@pytest.mark.parametrize("norb,bl", [(40, 20), (70, 20), (128, 20), (100, 64), (70, 13)])
def test_header_packs_strings_into_64_bit_words(shim_io, tmp_path, norb, bl):
    """The drop-in header's pack (the semantics of change_bitlength, bit_manipulation.h:863-893):
    bit_length-bit strings of norb orbitals become one 64-bit word each up to 64 orbitals and
    two (low word first) up to 128; more are refused."""
    rng = np.random.default_rng(norb + bl)
    masks = [sum(1 << int(p) for p in rng.choice(norb, int(rng.integers(1, min(norb, 20))), replace=False))
             for _ in range(50)]
    nw = (norb + bl - 1) // bl
    words = np.array([(m >> (bl * w)) & ((1 << bl) - 1) for m in masks for w in range(nw)], np.uint64)
    words.tofile(tmp_path / "s.u64")
    out = str(tmp_path / "p")
    r = subprocess.run([shim_io, "pack", str(tmp_path / "s.u64"), str(norb), str(bl), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    W = 1 if norb <= 64 else 2
    assert json.loads(r.stdout.strip().splitlines()[-1])["words"] == W
    want = [(m >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for m in masks for k in range(W)]
    assert np.array_equal(np.fromfile(out + ".packed.u64", np.uint64), np.array(want, np.uint64))


def test_header_refuses_more_than_128_orbitals(shim_io, tmp_path):
    np.zeros(7, np.uint64).tofile(tmp_path / "s.u64")  # one 129-orbital string of 20-bit words
    r = subprocess.run([shim_io, "pack", str(tmp_path / "s.u64"), "129", "20", str(tmp_path / "p")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "at most 128 orbitals" in r.stdout
