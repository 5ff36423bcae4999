"""Text datasets in the reference's format (one decimal key per line; range files as start/end line
pairs), parsed by the library's host code -- no GPU needed."""
import os

import numpy as np
import pytest

from harmonia_amd import HarmoniaError, read_keys_text, read_ranges_text
from harmonia_amd import workloads as W


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_keys_roundtrip_large(tmp_path):
    keys, _ = W.gen_keys(300000, 9)
    path = write(tmp_path, "insert.txt", "".join(f"{int(k)}\n" for k in keys))
    got = read_keys_text(path)
    assert got.dtype == np.uint64 and np.array_equal(got, keys)


def test_keys_edge_forms(tmp_path):
    # what the reference's generators write (1 .. 2^63-2), plus blanks, '+', CRLF, a missing final
    # newline and empty lines, which sscanf-style readers accept too
    text = "1\n  42  \n+7\r\n\n9223372036854775806\n0\n9223372036854775807"
    got = read_keys_text(write(tmp_path, "k.txt", text))
    assert got.tolist() == [1, 42, 7, (1 << 63) - 2, 0, (1 << 63) - 1]
    assert read_keys_text(write(tmp_path, "empty.txt", "")).size == 0


@pytest.mark.parametrize("bad", ["-5", "12a", "9223372036854775808", "1 2", "x", "+"])
def test_keys_rejects_malformed(tmp_path, bad):
    path = write(tmp_path, "bad.txt", f"1\n2\n{bad}\n3\n")
    with pytest.raises(HarmoniaError, match=":3:"):
        read_keys_text(path)


def test_ranges(tmp_path):
    g = W.rng(4, 1)
    s = g.integers(0, 1 << 62, 5000, dtype=np.uint64)
    e = s + g.integers(0, 1 << 20, 5000, dtype=np.uint64)
    text = "".join(f"{int(a)}\n{int(b)}\n" for a, b in zip(s, e))
    a, b = read_ranges_text(write(tmp_path, "range.txt", text))
    assert np.array_equal(a, s) and np.array_equal(b, e)
    with pytest.raises(HarmoniaError, match="odd"):
        read_ranges_text(write(tmp_path, "odd.txt", "1\n2\n3\n"))


def test_cpp_mirror_reads_dataset(tmp_path):
    """The C++ mirror (include/harmonia.hpp) compiles against the C ABI, links the library and reads a
    dataset through it (host code only)."""
    import shutil
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if not shutil.which("g++"):
        pytest.skip("no g++")
    src = tmp_path / "m.cpp"
    src.write_text("\n".join([
        '#include "harmonia.hpp"',
        "#include <cstdio>",
        "int main(int c, char** v) {",
        "  auto k = harmonia::read_keys_text(v[1]);",
        "  unsigned long long s = 0;",
        "  for (auto x : k) s += x;",
        '  std::printf("%zu %llu\\n", k.size(), s);',
        "  return 0;",
        "}", ""]))
    exe = tmp_path / "m"
    lib = os.path.join(root, "harmonia_amd")
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(root, "include"), str(src), "-L", lib,
                           "-lharmonia_amd", f"-Wl,-rpath,{lib}", "-o", str(exe)])
    data = write(tmp_path, "k.txt", "5\n7\n11\n")
    out = subprocess.check_output([str(exe), data], text=True).split()
    assert out == ["3", "23"]
