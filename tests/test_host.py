"""Host side of the library and its Python mirror, without a GPU.

Covers the C ABI (every symbol include/dto.h declares is exported, the record
layout matches the header), the host-computed parts of the path against the
oracle (thresholds, ln n! table, FDR, empirical p-value, JSON bytes), the
input readers and population rules (error codes and messages), the Philox
generator against its published known answers, and that every compute entry
point fails loudly when no device is present.
"""
import ctypes
import json
import math
import os
import re
import subprocess
import textwrap

import numpy as np
import pytest

from conftest import ROOT
from helpers import readme_lists

HEADER = os.path.join(ROOT, "include", "dto.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dto_[a-z0-9_]+)\s*\(", src)))


# ---- the ABI ---------------------------------------------------------------------

def test_every_declared_symbol_is_exported(dto):
    L = dto._lib.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(dto._lib.SIGNATURES), set(names) ^ set(dto._lib.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", dto._lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (dto_\w+)", out))
    assert set(names) <= exported


def test_struct_layout_matches_header(dto, tmp_path):
    """Compiles a probe against include/dto.h and compares sizes and offsets
    with the ctypes mirror."""
    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent("""
        #include <stddef.h>
        #include <stdio.h>
        #include "dto.h"
        #define F(T, m) printf("%s.%s %zu\\n", #T, #m, offsetof(T, m))
        int main(void) {
          printf("dto_record %zu\\n", sizeof(dto_record));
          printf("dto_result %zu\\n", sizeof(dto_result));
          printf("dto_job_info %zu\\n", sizeof(dto_job_info));
          printf("dto_server_config %zu\\n", sizeof(dto_server_config));
          F(dto_record, rank1); F(dto_record, rank2); F(dto_record, set1_len); F(dto_record, set2_len);
          F(dto_record, population_size); F(dto_record, intersection_size); F(dto_record, pvalue);
          F(dto_record, permuted); F(dto_record, idx1); F(dto_record, idx2);
          F(dto_result, unpermuted); F(dto_result, empirical_pvalue); F(dto_result, fdr);
          F(dto_result, n_permutations); F(dto_result, n_le);
          return 0;
        }
    """))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    m = dto._lib
    assert int(got["dto_record"]) == ctypes.sizeof(m.Record) == 72
    assert int(got["dto_result"]) == ctypes.sizeof(m.Result)
    assert int(got["dto_job_info"]) == ctypes.sizeof(m.JobInfo)
    assert int(got["dto_server_config"]) == ctypes.sizeof(m.ServerConfig)
    for k, v in got.items():
        if "." in k:
            t, f = k.split(".")
            cls = {"dto_record": m.Record, "dto_result": m.Result}[t]
            assert getattr(cls, f).offset == int(v), k
    for f in m.RECORD_DTYPE.names:
        assert m.RECORD_DTYPE.fields[f][1] == getattr(m.Record, f).offset


# ---- host-computed parts against the oracle ----------------------------------------

def test_thresholds_match_oracle(dto, O):
    from dual_threshold_optimization_amd.collections import generate_thresholds
    for m in list(range(0, 300)) + [469, 6000, 19999, 20000, 22000, 30000, 65535, 1_000_000, 2**32 - 1]:
        assert list(generate_thresholds(m)) == O.generate_thresholds(m), m


def test_ln_factorial_table_bitwise(dto, O):
    n = 30000
    out = np.empty(n + 1, dtype=np.float64)
    dto._lib.check(dto._lib.lib().dto_ln_factorial_table(n, dto._lib.dblp(out)))
    want = np.array([O.ln_factorial(x) for x in range(n + 1)])
    assert np.array_equal(out.view(np.uint64), want.view(np.uint64))


def test_fdr_bitwise(dto, O):
    from dual_threshold_optimization_amd.stat_operations import fdr
    rng = np.random.default_rng(5)
    for _ in range(2000):
        N = int(rng.integers(1, 40000))
        l1, l2 = int(rng.integers(0, N + 1)), int(rng.integers(0, N + 1))
        k = int(rng.integers(0, min(l1, l2) + 1))
        s = float(rng.choice([0.8, 0.5, 1.0, 0.33]))
        assert fdr(l1, l2, k, N, s) == O.fdr(l1, l2, k, N, s)
    assert abs(fdr(200, 150, 50, 1000) - 0.240625) < 1e-4  # fdr.rs:19-27
    with pytest.raises(ValueError, match="Sensitivity"):
        fdr(1, 1, 1, 1, 0.0)


def _record(rank1, rank2, s1, s2, N, k, p, permuted):
    from dual_threshold_optimization_amd._lib import Record
    r = Record()
    r.rank1, r.rank2, r.set1_len, r.set2_len, r.population_size, r.intersection_size = rank1, rank2, s1, s2, N, k
    r.pvalue, r.permuted = p, permuted
    return r


def _result_json(dto, unperm, perm_ps):
    m = dto._lib
    perm = (m.Record * max(len(perm_ps), 1))(*[_record(1, 1, 1, 1, 1, 1, p, 1) for p in perm_ps])
    res = m.Result()
    m.check(m.lib().dto_empirical_pvalue(ctypes.byref(unperm), ctypes.cast(perm, ctypes.c_void_p),
                                         len(perm_ps), ctypes.byref(res)))
    return res, dto.stat_operations.result_json(res)


def test_empirical_pvalue_json_example(dto):
    # src/stat_operations/empirical_pvalue.rs:40-107
    u = _record(1, 2, 50, 40, 100, 10, 0.05, 0)
    res, text = _result_json(dto, u, [0.10, 0.03, 0.07])
    assert res.n_permutations == 3 and res.n_le == 1
    assert json.loads(text) == {"rank1": 1, "rank2": 2, "set1_len": 50, "set2_len": 40, "population_size": 100,
                                "unpermuted_intersection_size": 10, "unpermuted_pvalue": 0.05,
                                "empirical_pvalue": 0.3333333333333333, "fdr": 1.03125}
    assert text == textwrap.dedent("""\
        {
          "empirical_pvalue": 0.3333333333333333,
          "fdr": 1.03125,
          "population_size": 100,
          "rank1": 1,
          "rank2": 2,
          "set1_len": 50,
          "set2_len": 40,
          "unpermuted_intersection_size": 10,
          "unpermuted_pvalue": 0.05
        }""")
    # no permutations: empirical p-value 1.0 (empirical_pvalue.rs:146-158)
    res, text = _result_json(dto, u, [])
    assert res.empirical_pvalue == 1.0 and '"empirical_pvalue": 1.0,' in text


def test_readme_json_bytes(dto):
    # README.md:172-186 (the output file of the README run, byte for byte)
    u = _record(24, 13, 24, 13, 30, 12, 0.15632183908046102, 0)
    res, text = _result_json(dto, u, [0.1, 0.01, 0.02, 0.04, 0.5])
    assert text == textwrap.dedent("""\
        {
          "empirical_pvalue": 0.8,
          "fdr": 0.0,
          "population_size": 30,
          "rank1": 24,
          "rank2": 13,
          "set1_len": 24,
          "set2_len": 13,
          "unpermuted_intersection_size": 12,
          "unpermuted_pvalue": 0.15632183908046102
        }""")


def ryu_layout(v):
    """serde_json's float layout (ryu crate, pretty.rs) on Python's shortest
    round-trip digits."""
    if v == 0:
        return "-0.0" if math.copysign(1, v) < 0 else "0.0"
    sign = "-" if v < 0 else ""
    m, e = ("%r" % abs(v)).lower().partition("e")[::2] if "e" in repr(abs(v)) else (repr(abs(v)), "0")
    exp = int(e)
    ip, _, fp = m.partition(".")
    digits = (ip + fp).lstrip("0")
    lead = len(ip.lstrip("0")) if ip.strip("0") else -(len(fp) - len(fp.lstrip("0")))
    kk = lead + exp
    digits = digits.rstrip("0") or "0"
    n = len(digits)
    k = kk - n
    if 0 <= k and kk <= 16:
        s = digits + "0" * k + ".0"
    elif 0 < kk <= 16:
        s = digits[:kk] + "." + digits[kk:]
    elif -5 < kk <= 0:
        s = "0." + "0" * (-kk) + digits
    elif n == 1:
        s = digits + "e" + str(kk - 1)
    else:
        s = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + s


def test_float_format_matches_serde_json(dto):
    def fmt(x):
        u = _record(1, 1, 1, 1, 1, 1, x, 0)
        _, text = _result_json(dto, u, [])
        return re.search(r'"unpermuted_pvalue": (.*)\n', text).group(1)
    # literal cases: serde_json / ryu output
    for x, s in [(1.0, "1.0"), (0.8, "0.8"), (0.0, "0.0"), (1.5e-7, "1.5e-7"), (1e-5, "0.00001"), (1e-6, "1e-6"),
                 (1e16, "1e16"), (1e15, "1000000000000000.0"), (123.456, "123.456"), (5e-324, "5e-324"),
                 (0.30000000000000004, "0.30000000000000004"), (1.7976931348623157e308, "1.7976931348623157e308"),
                 (2.5e-300, "2.5e-300"), (0.15632183908046102, "0.15632183908046102"),
                 (0.00044068070222441115, "0.00044068070222441115"), (float("nan"), "null")]:
        assert fmt(x) == s, (x, fmt(x), s)
    rng = np.random.default_rng(3)
    xs = np.concatenate([10.0 ** rng.uniform(-320, 20, 3000), rng.uniform(0, 1, 1000),
                         np.floor(rng.uniform(0, 1e17, 200))])
    for x in xs:
        assert fmt(float(x)) == ryu_layout(float(x)), float(x)


# ---- Philox generator ---------------------------------------------------------------

def test_philox_known_answers(dto):
    """Philox4x32-10 (Salmon et al., SC'11), published known-answer vectors."""
    L = dto._lib.lib()
    for ctr, key, want in [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]:
        c = np.array(ctr, dtype=np.uint32)
        k = np.array(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        assert L.dto_internal_philox(dto._lib.u32p(c), dto._lib.u32p(k), dto._lib.u32p(out)) == 0
        assert list(out) == want


def test_philox_shuffle_is_a_permutation(dto):
    from dual_threshold_optimization_amd.collections import philox_permutation
    for n in [0, 1, 2, 3, 5, 64, 1000, 4096, 4097, 20000, 65535]:
        p = philox_permutation(7, 0, 1, 1, n)
        assert sorted(p.tolist()) == list(range(n))
        assert np.array_equal(p, philox_permutation(7, 0, 1, 1, n))
    a = philox_permutation(7, 0, 1, 1, 20000)
    for other in [philox_permutation(8, 0, 1, 1, 20000), philox_permutation(7, 1, 1, 1, 20000),
                  philox_permutation(7, 0, 2, 1, 20000), philox_permutation(7, 0, 1, 2, 20000)]:
        assert np.mean(a == other) < 0.01
    # a uniform shuffle fixes ~1 point and moves a position anywhere
    assert np.sum(a == np.arange(20000)) < 12
    firsts = np.array([philox_permutation(s, 0, 0, 1, 10)[0] for s in range(2000)])
    counts = np.bincount(firsts, minlength=10)
    assert counts.min() > 140 and counts.max() < 260


def test_philox_shuffle_uniformity(dto):
    """Position-by-value counts of many shuffles against a uniform shuffle:
    chi-square of the full table at n = 30 (power-of-two floor domain), and
    of a 50 x 50 binned table at n = 20,000 (mixed-radix domain, 1 % above n);
    adjacent positions correlate as a uniform shuffle's do, -1 / (n - 1)."""
    from dual_threshold_optimization_amd.collections import philox_permutation

    def z_score(n, S, B):
        M = np.zeros((B, B))
        for s in range(S):
            p = philox_permutation(1000 + s, 5, s, 2, n).astype(np.int64)
            np.add.at(M, ((np.arange(n) * B) // n, (p * B) // n), 1)
        exp = M.sum() / (B * B)
        chi = ((M - exp) ** 2 / exp).sum()
        dof = (B - 1) ** 2
        # each cell is a sum of S hypergeometric draws (n/B positions, n/B
        # values), so the statistic's mean is dof * n / (n - 1), not dof
        return (chi - dof * n / (n - 1)) / np.sqrt(2 * dof)
    assert abs(z_score(30, 12000, 30)) < 4
    assert abs(z_score(20000, 120, 50)) < 4
    adj = [np.corrcoef(p[:-1], p[1:])[0, 1] for p in
           (philox_permutation(s, 0, s, 1, 1000).astype(np.float64) for s in range(1500))]
    assert abs(np.mean(adj) + 1 / 999) < 4 * np.std(adj) / np.sqrt(len(adj))
    # the joint law at n = 4: all 24 orders equally frequent (multinomial chi-square, 23 dof)
    counts = {}
    for s in range(24000):
        key = tuple(philox_permutation(5000 + s, 1, s, 1, 4).tolist())
        counts[key] = counts.get(key, 0) + 1
    assert len(counts) == 24
    chi = sum((c - 1000) ** 2 / 1000 for c in counts.values())
    assert abs(chi - 23) / np.sqrt(46) < 4


# ---- inputs, population rules and errors ---------------------------------------------

def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_job_errors_before_the_device(dto, tmp_path):
    from dual_threshold_optimization_amd._lib import (DTO_ERR_IO, DTO_ERR_PARSE, DTO_ERR_POPULATION, DtoError)
    g = ["gene1", "gene2", "gene3"]
    r = [1, 2, 3]
    with pytest.raises(DtoError) as e:
        dto.Job.from_lists(g, r, ["gene1", "gene2", "gene4"], r)
    assert e.value.code == DTO_ERR_POPULATION and "identical genes" in e.value.message
    with pytest.raises(DtoError) as e:
        dto.Job.from_lists(g, r, g, r, background=["gene1", "gene2"])
    assert e.value.code == DTO_ERR_POPULATION and "first ranked feature list" in e.value.message
    with pytest.raises(DtoError) as e:
        dto.Job.from_lists(g, r, ["gene1", "gene2", "gene4"], r, background=g)
    assert e.value.code == DTO_ERR_POPULATION and "second ranked feature list" in e.value.message
    good = _write(tmp_path, "a.csv", "gene1,1\ngene2,2\n")
    with pytest.raises(DtoError) as e:
        dto.Job.from_files(str(tmp_path / "missing.csv"), good)
    assert e.value.code == DTO_ERR_IO
    bad = _write(tmp_path, "b.csv", "gene1,1\ngene2,2,3\n")
    with pytest.raises(DtoError) as e:
        dto.Job.from_files(good, bad)
    assert e.value.code == DTO_ERR_PARSE and "Invalid format in file %s at line 2" % bad in e.value.message
    for rank in ["-1", "1.5", "", "x", "18446744073709551616"]:
        bad = _write(tmp_path, "c.csv", "gene1,1\ngene2,%s\n" % rank)
        with pytest.raises(DtoError) as e:
            dto.Job.from_files(good, bad)
        assert e.value.code == DTO_ERR_PARSE and "Invalid rank value" in e.value.message, rank


def test_compute_fails_loudly_without_device(dto):
    """With no GPU every compute entry point returns DTO_ERR_NO_DEVICE: there
    is no CPU fallback."""
    from dual_threshold_optimization_amd._lib import DTO_ERR_NO_DEVICE, DtoError
    if dto._lib.device_count() > 0:
        pytest.skip("a device is present")
    i1, r1, i2, r2 = readme_lists()
    with pytest.raises(DtoError) as e:
        dto.Job.from_lists(i1, r1, i2, r2)
    assert e.value.code == DTO_ERR_NO_DEVICE
    with pytest.raises(DtoError) as e:
        dto.stat_operations.hypergeometric_pvalue(1000, 50, 60, 10)
    assert e.value.code == DTO_ERR_NO_DEVICE
    with pytest.raises(DtoError) as e:
        dto.Server()
    assert e.value.code == DTO_ERR_NO_DEVICE


def test_readers_match_reference_semantics(dto, tmp_path):
    from dual_threshold_optimization_amd.read import read_feature_list_from_file, read_ranked_feature_list_from_csv
    # trim, '+' sign, CRLF, rank `as u32` wrap, stable order of ties
    p = _write(tmp_path, "r.csv", " b , 2\r\na,+1\nc,4294967298\nd, 2 \n")
    rl = read_ranked_feature_list_from_csv(p)
    assert rl.ids() == ["a", "b", "c", "d"] and list(rl.ranks()) == [1, 2, 2, 2]
    assert list(rl.thresholds()) == [1, 2]
    with pytest.raises(ValueError, match="Invalid format in file"):
        read_ranked_feature_list_from_csv(_write(tmp_path, "s.csv", "a,1\n\n"))
    with pytest.raises(ValueError, match="Invalid rank value"):
        read_ranked_feature_list_from_csv(_write(tmp_path, "t.csv", "a,-3\n"))
    # background: one id per line, trimmed; blank lines are "" ids and count
    bg = read_feature_list_from_file(_write(tmp_path, "bg.txt", "g1\n g2 \n\ng3"))
    assert bg.ids() == ["g1", "g2", "", "g3"] and len(bg) == 4


def test_python_mirror_collections(dto):
    from dual_threshold_optimization_amd.collections import (FeatureList, PermutedRankedFeatureList,
                                                             RankedFeatureList)
    from dual_threshold_optimization_amd.dto import compute_population_size
    from dual_threshold_optimization_amd.stat_operations import intersect_genes
    # ranked.rs:674-696: iteration in rank order
    rl = RankedFeatureList.from_(FeatureList(["gene1", "gene2", "gene3"]), [3, 2, 1])
    assert rl.ids() == ["gene3", "gene2", "gene1"] and list(rl.ranks()) == [1, 2, 3]
    assert [g.id() for g in rl.get_feature_set_by_threshold(2)] == ["gene3", "gene2"]
    with pytest.raises(ValueError):
        RankedFeatureList.from_(FeatureList(["a", "b"]), [1])
    # permuted.rs:90-101: ranks stay at positions, features move
    pr = PermutedRankedFeatureList(rl, indices=[2, 0, 1])
    assert [g.id() for g in pr.get_feature_set_by_threshold(2)] == ["gene1", "gene3"]
    # permuted.rs:110-136: ten features below the threshold keep their count
    big = RankedFeatureList.from_(FeatureList(["g%d" % i for i in range(20)]), list(range(1, 21)))
    assert len(PermutedRankedFeatureList(big, seed=3).get_feature_set_by_threshold(10)) == 10
    # intersect_genes.rs:25-36 and the duplicate asymmetry
    assert intersect_genes(["gene1", "gene2", "gene3"], ["gene2", "gene3", "gene4"]) == 2
    assert intersect_genes(["a"], ["a", "a"]) == 2 and intersect_genes(["a", "a"], ["a"]) == 1
    # compute_population_size.rs:111-258
    g = FeatureList(["gene1", "gene2", "gene3"])
    a = RankedFeatureList.from_(g, [1, 2, 3])
    b = RankedFeatureList.from_(FeatureList(["gene1", "gene2", "gene4"]), [1, 2, 3])
    assert compute_population_size(a, a) == 3
    assert compute_population_size(a, a, FeatureList(["gene1", "gene2", "gene3", "gene4"])) == 4
    with pytest.raises(ValueError, match="identical genes"):
        compute_population_size(a, b)
    with pytest.raises(ValueError, match="first ranked"):
        compute_population_size(a, a, FeatureList(["gene1", "gene2"]))
    with pytest.raises(ValueError, match="second ranked"):
        compute_population_size(a, b, g)


# ---- the CLI ------------------------------------------------------------------------------

def _cli(dto, *args):
    return subprocess.run([dto._lib.CLI_PATH] + list(args), capture_output=True, text=True, timeout=60)


def test_cli_arguments(dto, tmp_path):
    if not os.path.exists(dto._lib.CLI_PATH):
        pytest.fail("dto_amd is not built")
    r = _cli(dto, "--help")
    assert r.returncode == 0 and "--ranked-list1" in r.stdout and "--permutations" in r.stdout
    r = _cli(dto, "-1", "x.csv")
    assert r.returncode == 2 and "--ranked-list2 <FILE>" in r.stderr
    r = _cli(dto, "-1", "a", "-2", "b", "-p", "ten")
    assert r.returncode == 2
    r = _cli(dto, "-1", "a", "-2", "b", "--frobnicate")
    assert r.returncode == 2 and "unexpected argument" in r.stderr
    # a missing file panics (exit status 101, as a Rust panic)
    r = _cli(dto, "-1", str(tmp_path / "nope.csv"), "-2", str(tmp_path / "nope2.csv"), "-p", "3")
    assert r.returncode == 101 and "panicked" in r.stderr
    assert "Ranked list 1: %s" % (tmp_path / "nope.csv") in r.stderr and "Permutations: 3" in r.stderr
    bad = _write(tmp_path, "bad.csv", "a,1,2\n")
    r = _cli(dto, "-1", bad, "-2", bad)
    assert r.returncode == 101 and "Invalid format in file" in r.stderr
    r = _cli(dto, "-1", bad, "-2", bad, "-t", "0")
    assert "Warning: Number of threads cannot be 0. Setting threads to 1." in r.stderr


def test_cli_without_device_fails_loudly(dto):
    if dto._lib.device_count() > 0:
        pytest.skip("a device is present")
    g = os.path.join(ROOT, "tests", "golden")
    r = _cli(dto, "-1", os.path.join(g, "ranklist1.csv"), "-2", os.path.join(g, "ranklist2.csv"), "-p", "5")
    assert r.returncode == 101 and "no HIP device" in r.stderr and r.stdout == ""


def _write_bytes(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def test_readers_unicode_trim_and_invalid_utf8(dto, tmp_path):
    """str::trim strips Unicode White_Space (read_ranked_feature_list_from_csv.rs:62-63,
    read_feature_list_from_file.rs:50); a line that is not valid UTF-8 makes
    lines() fail and the reader panic (read_ranked_feature_list_from_csv.rs:57).
    The library's readers and the Python mirror agree with both rules."""
    from dual_threshold_optimization_amd import _lib
    from dual_threshold_optimization_amd.read import read_feature_list_from_file, read_ranked_feature_list_from_csv
    ws = "  　\u0085    \t\x0b"
    text = ("%sgA%s,%s3%s\r\n" % (ws, ws, ws, ws)) + "gB​,+1\n" + "\x1cgC\x1f,2\n"
    p = _write_bytes(tmp_path, "u.csv", text.encode())
    ids, ranks = _lib.read_ranked_csv_native(p)
    # U+200B (zero width space) and U+001C/U+001F are not White_Space: kept
    assert ids == ["gA".encode(), "gB​".encode(), "\x1cgC\x1f".encode()]
    assert list(ranks) == [3, 1, 2]
    rl = read_ranked_feature_list_from_csv(p)
    assert rl.ids() == ["gB​", "\x1cgC\x1f", "gA"] and list(rl.ranks()) == [1, 2, 3]
    bgp = _write_bytes(tmp_path, "bg.txt", ("%sx%s\n \ny\x1c\n" % (ws, ws)).encode())
    assert _lib.read_feature_lines_native(bgp) == [b"x", b"", "y\x1c".encode()]
    assert read_feature_list_from_file(bgp).ids() == ["x", "", "y\x1c"]
    bad_lines = [b"a\xff,1\n", b"a\xc0\x80,1\n", b"a\xed\xa0\x80,1\n", b"a\xf4\x90\x80\x80,1\n", b"a\xe2\x82,1\n"]
    for k, bad in enumerate(bad_lines):
        p = _write_bytes(tmp_path, "bad%d.csv" % k, b"ok,1\n" + bad)
        with pytest.raises(_lib.DtoError) as e:
            _lib.read_ranked_csv_native(p)
        assert e.value.code == _lib.DTO_ERR_PARSE and "valid UTF-8" in e.value.message
        with pytest.raises(ValueError, match="valid UTF-8"):
            read_ranked_feature_list_from_csv(p)
        pb = _write_bytes(tmp_path, "badbg%d.txt" % k, b"ok\n" + bad)
        with pytest.raises(_lib.DtoError) as e:
            _lib.read_feature_lines_native(pb)
        assert e.value.code == _lib.DTO_ERR_PARSE
        with pytest.raises(ValueError, match="valid UTF-8"):
            read_feature_list_from_file(pb)
    # well-formed 2-, 3- and 4-byte sequences are accepted
    p = _write_bytes(tmp_path, "ok.csv", "é,1\n€,2\n\U0001f600,3\n".encode())
    assert _lib.read_ranked_csv_native(p)[0] == ["é".encode(), "€".encode(), "\U0001f600".encode()]
    # lines are read in order: a format error before the bad line wins, and after it loses
    p = _write_bytes(tmp_path, "o1.csv", b"a,1,2\nb\xff,1\n")
    with pytest.raises(_lib.DtoError, match="Invalid format"):
        _lib.read_ranked_csv_native(p)
    with pytest.raises(ValueError, match="Invalid format"):
        read_ranked_feature_list_from_csv(p)
    p = _write_bytes(tmp_path, "o2.csv", b"b\xff,1\na,1,2\n")
    with pytest.raises(_lib.DtoError, match="valid UTF-8"):
        _lib.read_ranked_csv_native(p)
    with pytest.raises(ValueError, match="valid UTF-8"):
        read_ranked_feature_list_from_csv(p)


def test_fdr_sensitivity_error_path(dto):
    """fdr.rs:36-39 panics for sensitivity <= 0: the ABI reports it as an error."""
    from dual_threshold_optimization_amd import _lib
    from dual_threshold_optimization_amd.stat_operations import fdr
    import ctypes
    import math
    out = ctypes.c_double()
    for s in (0.0, -0.5, -0.0):
        assert _lib.lib().dto_fdr_checked(200, 150, 50, 1000, s, ctypes.byref(out)) == _lib.DTO_ERR_INVALID_ARGUMENT
        assert _lib.lib().dto_last_error() == b"Sensitivity must be greater than 0."
        assert math.isnan(_lib.lib().dto_fdr(200, 150, 50, 1000, s))
        with pytest.raises(ValueError, match="Sensitivity must be greater than 0."):
            fdr(200, 150, 50, 1000, s)
    assert _lib.lib().dto_fdr_checked(200, 150, 50, 1000, 0.8, ctypes.byref(out)) == 0
    assert abs(out.value - 0.240625) < 1e-4 and out.value == fdr(200, 150, 50, 1000)
