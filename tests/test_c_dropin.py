"""The two f5c drop-ins called from plain C (tests/native/c_dropin.c), as the
C shim added to f5c calls them: the header compiles as C11 with the f5c
type layouts checked, the program links against the in-tree libabea, and on
the GPU both entry points return the reference's pairs for the golden reads
(GPU/f5cmisc.h:48-50, :99-102)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

EXE = os.path.join(ROOT, "tests", "native", "_build", "c_dropin")


def test_header_compiles_as_c_and_program_links(tmp_path):
    src = os.path.join(ROOT, "tests", "native", "c_dropin.c")
    lib = os.path.join(ROOT, "abea-opencl_amd", "lib", "libabea.so")
    if not os.path.exists(lib):
        pytest.skip("libabea is not built")
    exe = str(tmp_path / "c_dropin")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic",
                    "-I", os.path.join(ROOT, "include"), "-o", exe, src,
                    "-L", os.path.dirname(lib), "-labea", "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True)
    assert os.path.getsize(exe) > 0


def _write_fixture(path, b):
    aos = b.events_aos()
    model = np.ascontiguousarray(b.model, np.float32)
    scal = np.ascontiguousarray(b.scalings(), np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", b.n_reads))
        f.write(model.tobytes())
        for i in range(b.n_reads):
            seq = b.read_seq(i)
            ev = np.ascontiguousarray(aos[b.ev_ptr[i]:b.ev_ptr[i] + b.n_events[i]])
            f.write(struct.pack("<ii", len(seq), int(b.n_events[i])))
            f.write(scal[i].tobytes())
            f.write(seq)
            f.write(ev.tobytes())


def _read_pairs(buf, off):
    (n,) = struct.unpack_from("<i", buf, off)
    off += 4
    pairs = np.frombuffer(buf, np.int32, 2 * n, off).reshape(-1, 2)
    return pairs, off + 8 * n


@pytest.mark.gpu
def test_c_program_returns_golden_pairs(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from abea_amd import load_golden_subset
    exe = EXE
    if not os.path.exists(exe):
        import __graft_entry__
        exe = __graft_entry__.build_c_dropin()
    b, golden, _ = load_golden_subset(GOLDEN)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_fixture(fin, b)
    print(f"[c_dropin] {b.n_reads} golden reads", flush=True)
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    buf = fout.read_bytes()
    off = 0
    for form in ("align", "align_db"):
        for i in range(b.n_reads):
            pairs, off = _read_pairs(buf, off)
            assert np.array_equal(pairs, golden[i]), (form, i)
    assert off == len(buf)
