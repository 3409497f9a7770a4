"""CPU tests of the tuning cache (the batched pass's persisted tuner choices).

Modelled on the reference's wisdom tests: the file is found through a directory resolved on every
call, and a missing, corrupt, or unwritable cache never fails the caller -- it only means the
next pass is tuned again (src/plans.jl:133-257, where FFTW wisdom is handled the same way).
"""
import os
import stat

import pytest

from autorift import batch

DEV = "gfx950-256cu"


@pytest.fixture
def cache_dir(tmp_path, monkeypatch):
    d = tmp_path / "tune"
    monkeypatch.setenv("AR_TUNE_CACHE_DIR", str(d))
    monkeypatch.delenv("AR_TUNE_CACHE", raising=False)
    return d


def _tmp_left(d):
    return [p for p in os.listdir(d) if p.endswith(".tmp")]


def test_path_follows_the_environment(cache_dir, tmp_path, monkeypatch):
    p = batch.tune_cache_path(DEV)
    assert p is not None and os.path.dirname(p) == str(cache_dir)
    assert cache_dir.is_dir()                       # made on demand
    assert "gfx950-256cu" in os.path.basename(p)
    # resolved on every call: a redirect takes effect at once
    other = tmp_path / "elsewhere"
    monkeypatch.setenv("AR_TUNE_CACHE_DIR", str(other))
    assert os.path.dirname(batch.tune_cache_path(DEV)) == str(other)
    # characters that do not belong in a file name are replaced
    q = batch.tune_cache_path("gfx950 / odd:name")
    assert "/" not in os.path.basename(q) and " " not in os.path.basename(q)
    monkeypatch.setenv("AR_TUNE_CACHE", "off")
    assert batch.tune_cache_path(DEV) is None
    assert batch.tune_cache_record(DEV, "k", 1) is False
    assert batch.tune_cache_lookup(DEV, "k") is None


def test_record_and_lookup(cache_dir):
    assert batch.tune_cache_lookup(DEV, "batch c32x32 r40x40 n244 e0") is None   # no file yet
    assert batch.tune_cache_record(DEV, "batch c32x32 r40x40 n244 e0", 16 | (28 << 8))
    assert batch.tune_cache_record(DEV, "batch c64x64 r25x25 n10 e1", 0)
    assert batch.tune_cache_lookup(DEV, "batch c32x32 r40x40 n244 e0") == 16 | (28 << 8)
    assert batch.tune_cache_lookup(DEV, "batch c64x64 r25x25 n10 e1") == 0
    assert batch.tune_cache_record(DEV, "batch c32x32 r40x40 n244 e0", 24)      # overwrite
    assert batch.tune_cache_lookup(DEV, "batch c32x32 r40x40 n244 e0") == 24
    # another device's file is separate
    assert batch.tune_cache_lookup("gfx950-128cu", "batch c32x32 r40x40 n244 e0") is None
    assert _tmp_left(cache_dir) == []
    # keys that could not be stored as one line are refused, not written
    assert batch.tune_cache_record(DEV, "bad\tkey", 1) is False
    assert batch.tune_cache_record(DEV, "bad\nkey", 1) is False
    assert batch.tune_cache_record(DEV, "", 1) is False


def test_corrupt_file_reads_as_empty_and_is_replaced(cache_dir):
    path = batch.tune_cache_path(DEV)
    with open(path, "wb") as f:
        f.write(os.urandom(3000) + b"\n\x00\xff garbage")
    assert batch.tune_cache_lookup(DEV, "k") is None
    assert batch.tune_cache_record(DEV, "k", 7)
    assert batch.tune_cache_lookup(DEV, "k") == 7
    text = open(path).read().splitlines()
    assert text[0] == "autorift_hip tune cache 1" and text[1:] == ["k\t7"]


def test_damaged_lines_are_skipped(cache_dir):
    path = batch.tune_cache_path(DEV)
    long_key = "x" * 5000
    with open(path, "w") as f:
        f.write("autorift_hip tune cache 1\n")
        f.write("good\t12\n")
        f.write("no tab here\n")
        f.write("\t5\n")                         # empty key
        f.write("empty\t\n")
        f.write("notanumber\t1x\n")
        f.write("huge\t99999999999999\n")        # outside int32
        f.write(long_key + "\t3\n")              # longer than a line buffer
        f.write("after\t-4\n")
        f.write("truncated\t8")                   # no final newline (a cut-off write)
    assert batch.tune_cache_lookup(DEV, "good") == 12
    assert batch.tune_cache_lookup(DEV, "after") == -4
    assert batch.tune_cache_lookup(DEV, "truncated") == 8
    for k in ("no tab here", "empty", "notanumber", "huge", long_key):
        assert batch.tune_cache_lookup(DEV, k) is None
    # a record keeps what parsed and drops the rest
    assert batch.tune_cache_record(DEV, "new", 1)
    lines = open(path).read().splitlines()
    assert sorted(lines[1:]) == sorted(["good\t12", "after\t-4", "truncated\t8", "new\t1"])


def test_wrong_header_is_treated_as_empty(cache_dir):
    path = batch.tune_cache_path(DEV)
    with open(path, "w") as f:
        f.write("some other format 9\nk\t3\n")
    assert batch.tune_cache_lookup(DEV, "k") is None


def test_unusable_directory_never_fails(tmp_path, monkeypatch):
    # the cache directory would have to be made below a regular file: no cache, no error
    blocker = tmp_path / "file"
    blocker.write_text("x")
    monkeypatch.setenv("AR_TUNE_CACHE_DIR", str(blocker / "tune"))
    monkeypatch.delenv("AR_TUNE_CACHE", raising=False)
    assert batch.tune_cache_path(DEV) is None
    assert batch.tune_cache_record(DEV, "k", 1) is False
    assert batch.tune_cache_lookup(DEV, "k") is None


def test_unwritable_cache_file_is_dropped(cache_dir):
    # the cache file's name is taken by a directory: the rename fails, the write is dropped and
    # no temporary is left behind
    path = batch.tune_cache_path(DEV)
    os.mkdir(path)
    os.mkdir(os.path.join(path, "child"))       # a non-empty directory cannot be renamed over
    assert batch.tune_cache_record(DEV, "k", 1) is False
    assert batch.tune_cache_lookup(DEV, "k") is None
    assert _tmp_left(cache_dir) == []


@pytest.mark.skipif(os.geteuid() == 0, reason="permission bits do not bind the superuser")
def test_read_only_directory_is_dropped(cache_dir):
    assert batch.tune_cache_record(DEV, "k", 1)
    os.chmod(cache_dir, stat.S_IRUSR | stat.S_IXUSR)
    try:
        assert batch.tune_cache_record(DEV, "k", 2) is False
        assert batch.tune_cache_lookup(DEV, "k") == 1      # the old file still reads
    finally:
        os.chmod(cache_dir, stat.S_IRWXU)
