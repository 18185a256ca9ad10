"""Training-set tools (tools/construct_train.in/construct_train_in.cpp and
expand222.cpp of the reference), host only.

* The converter against the reference's own converted files: the wurtzite
  Si and GaAs train.in (develop/wurzite/{Si,GaAs}) were written by variants
  of the converter from the DFT data files bundled here.  Those variants add
  an author-specific relabelling of the atom types (develop/wurzite/Si/
  construct_train_in.cpp:54-56,142-149,221-236: type 1 inside three z bands
  of the force cells, types {0, 1, 1, 0} in all but the last three energy
  cells, type 0 elsewhere in GaAs); the test applies that rule to the
  converter's output and then expects the reference file byte for byte.
* The energy shift and the kbar -> eV conversion against values fixed by
  arithmetic.
* The expander through invariants of a periodic crystal: a 2 x 2 x 2
  supercell has 8x the energy and virial and the same force on each image
  (FP64 oracle on both files).
"""
import os

import numpy as np
import pytest

from conftest import read_ga_out

BANDS = ((1.5, 4.5), (7.5, 10.5), (13.5, 17.0))
CELL_TYPES = (0, 1, 1, 0)


def unpack_dft(golden_dir, dest):
    """Write the bundled input.txt and DFT data files under dest."""
    text = open(os.path.join(golden_dir, "dft_inputs.txt")).read()
    name, buf = None, []
    files = {}
    for line in text.split("\n"):
        if line == "#list" or line.startswith("#file "):
            if name is not None:
                files[name] = "\n".join(buf) + "\n"
            name, buf = ("input.txt" if line == "#list" else line[6:]), []
        else:
            buf.append(line)
    files[name] = "\n".join(buf).rstrip("\n") + "\n"
    for n, t in files.items():
        p = os.path.join(dest, n)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            f.write(t)
    return os.path.join(dest, "input.txt")


def relabel(text, keep_file_type):
    lines = text.split("\n")
    nc, ncf = map(int, lines[0].split())
    na = [int(lines[1 + n]) for n in range(nc)]
    out, k = lines[:1 + nc], 1 + nc
    for n in range(nc):
        if n < ncf:
            out.append(lines[k])
            k += 1
            for _ in range(na[n]):
                t = lines[k].split(" ")
                typ = int(t[0]) if keep_file_type else 0
                if any(lo < float(t[3]) < hi for lo, hi in BANDS):
                    typ = 1
                t[0] = str(typ)
                out.append(" ".join(t))
                k += 1
        else:
            out += lines[k:k + 2]
            k += 2
            cell = na[n] // 48
            for a in range(na[n]):
                t = lines[k].split(" ")
                if n < nc - 3:
                    t[0] = str(CELL_TYPES[a % cell])
                elif not keep_file_type:
                    t[0] = "0"
                out.append(" ".join(t))
                k += 1
    return "\n".join(out + lines[k:])


@pytest.mark.parametrize("name,delta_e,keep", [("si_type2", 1.31684, True), ("gaas", 0.0, False)])
def test_construct_reproduces_reference_train_in(golden, tmp_path, name, delta_e, keep):
    import gpuga_amd as G
    d = golden(name)
    lst = unpack_dft(d, str(tmp_path))
    out = tmp_path / "train.in"
    G.construct_train_in(lst, out, cells=(4, 4, 3), delta_e=delta_e, stress_in_kbar=True)
    mine = out.read_text()
    ref = open(os.path.join(d, "train.in")).read().replace("\r\n", "\n")
    assert relabel(mine, keep) == ref


def _write(p, text):
    p.write_text(text)
    return str(p)


def test_construct_arithmetic(tmp_path):
    import gpuga_amd as G
    # energy cell: 2 atoms, E = -10 eV, every stress component 1602.18 kbar,
    # box 2 x 3 x 4 (volume 24 A^3); force cell copied as read
    _write(tmp_path / "e", "2\n-10\n1602.18 1602.18 1602.18 0 0 -1602.18\n"
           "2 0 0\n0 3 0\n0 0 4\n0 0.5 0.5 0.5\n1 1.5 1.5 1.5\n")
    _write(tmp_path / "f", "1\n5 0 0 0 5 0 0 0 5\n3 1.25 -2.5 0.125 0.5 -0.25 1e-3\n")
    lst = _write(tmp_path / "list", "1 1\nf\ne\n")
    G.construct_train_in(lst, tmp_path / "train.in", cells=(1, 1, 2), delta_e=1.0)
    lines = (tmp_path / "train.in").read_text().split("\n")
    assert lines[:3] == ["2 1", "1", "4"]
    assert lines[3] == "5 0 0 0 5 0 0 0 5 "
    assert lines[4] == "3 1.25 -2.5 0.125 0.5 -0.25 0.001"
    # ((E / 2) + 1) x 2 x 2 cells = -16; 1602.18 x 2 cells x 24 / 1602.18 = 48
    assert [float(x) for x in lines[5].split()] == [-16, 48, 48, 48, 0, 0, -48]
    assert lines[6] == "2 0 0 0 3 0 0 0 8 "  # c doubled
    assert lines[7:11] == ["0 0.5 0.5 0.5", "1 1.5 1.5 1.5",   # image 0
                           "0 0.5 0.5 4.5", "1 1.5 1.5 5.5"]   # image + c
    # stress already in eV: only the cell count scales the virial
    G.construct_train_in(lst, tmp_path / "t2", cells=(1, 1, 2), delta_e=0.0,
                         stress_in_kbar=False)
    e = [float(x) for x in (tmp_path / "t2").read_text().split("\n")[5].split()]
    assert e[0] == -20 and e[1] == pytest.approx(3204.36, rel=1e-7)
    # expand222's text: atom counts x 8, box x 2, the 8 images of an atom in
    # a row (c fastest), energy and virial x 8 with a space after each
    G.expand_train_in(tmp_path / "train.in", tmp_path / "x")
    x = (tmp_path / "x").read_text().split("\n")
    assert x[:4] == ["2 1", "8", "32", "10 0 0 0 10 0 0 0 10 "]
    assert x[4:7] == ["3 1.25 -2.5 0.125 0.5 -0.25 0.001", "3 1.25 -2.5 5.125 0.5 -0.25 0.001",
                      "3 1.25 2.5 0.125 0.5 -0.25 0.001"]
    assert x[11] == "3 6.25 2.5 5.125 0.5 -0.25 0.001"
    assert x[12:14] == ["-128 384 384 384 0 0 -384 ", "4 0 0 0 6 0 0 0 16 "]
    assert x[14:16] == ["0 0.5 0.5 0.5", "0 0.5 0.5 8.5"]


def test_construct_errors(tmp_path):
    import gpuga_amd as G
    lst = _write(tmp_path / "list", "1 0\nmissing\n")
    with pytest.raises(G.GpugaError, match="Failed to open"):
        G.construct_train_in(lst, tmp_path / "o")
    _write(tmp_path / "f", "2\n5 0 0 0 5 0 0 0 5\n0 1 1 1 0 0 0\n")  # one atom short
    lst = _write(tmp_path / "list", "1 0\nf\n")
    with pytest.raises(G.GpugaError, match="Reading error"):
        G.construct_train_in(lst, tmp_path / "o")


def test_expand_is_a_periodic_supercell(golden, tmp_path):
    """expand222 on a small converted set (two force and two energy cells):
    oracle energies and virials x 8, forces equal on the 8 images, neighbour
    counts equal (a wrong image offset overlaps atoms or leaves gaps, which
    changes the counts and the energy by orders of magnitude more)."""
    import gpuga_amd as G
    from oracle import oracle as O
    d = golden("si_type2")
    unpack_dft(d, str(tmp_path))
    names = open(tmp_path / "input.txt").read().split()[2:]
    lst = _write(tmp_path / "small", "2 2\n%s\n" % "\n".join(names[:2] + names[-2:]))
    one, two = tmp_path / "one", tmp_path / "two"
    one.mkdir()
    two.mkdir()
    for p in (one, two):
        (p / "potential.in").write_text(open(os.path.join(d, "potential.in")).read())
    G.construct_train_in(lst, one / "train.in", cells=(4, 4, 3), delta_e=1.31684)
    G.expand_train_in(one / "train.in", two / "train.in")
    a, b = O.Case(str(one)), O.Case(str(two))
    assert b.Nc == a.Nc == 4 and b.N == 8 * a.N
    params = read_ga_out(os.path.join(d, "ga.out"))[-1][2]
    fa, wa, ea = a.find_force(params)
    fb, wb, eb = b.find_force(params)
    Na, Nb = a.array("Na", 4), b.array("Na", 4)
    oa, ob = a.array("Na_sum", 4), b.array("Na_sum", 4)
    assert np.array_equal(Nb, 8 * Na)
    nna, nnb = a.neighbors()[0], b.neighbors()[0]
    for c in range(4):
        ia = np.arange(oa[c], oa[c] + Na[c])
        ib = ob[c] + 8 * np.arange(Na[c])   # images of an atom are consecutive
        # the images' coordinates are printed to 6 digits like the originals,
        # so coordinates above 10 A move by up to 5e-5 A: energies agree to
        # 1e-6; forces move by that displacement times the stiffness (2e-3
        # eV/A), and the virial by about 1e-3 eV per bond, which adds up to
        # about 0.1 eV over 6 000 bonds (diagonal components are ~70 eV)
        assert eb[ob[c]:ob[c] + Nb[c]].sum() == pytest.approx(8 * ea[ia].sum(), rel=1e-6)
        for k in range(6):
            assert wb[k * b.N + ob[c]:k * b.N + ob[c] + Nb[c]].sum() == pytest.approx(
                8 * wa[k * a.N + ia].sum(), abs=0.3)
        for img in range(8):
            assert np.array_equal(nnb[ib + img], nna[ia])
            for k in range(3):
                np.testing.assert_allclose(fb[k * b.N + ib + img], fa[k * a.N + ia], atol=2e-3)
    # references: force rows copied to every image, energy and virial x 8
    ra, rb = a.array("force_ref", 3 * a.N), b.array("force_ref", 3 * b.N)
    for k in range(3):
        assert np.array_equal(rb[k * b.N:k * b.N + 8 * Na[0]:8], ra[k * a.N:k * a.N + Na[0]])
    pa, pb = a.array("pe_ref", 4), b.array("pe_ref", 4)
    np.testing.assert_allclose(pb[2:], 8 * pa[2:], rtol=5e-6)  # printed to 6 digits


def test_cli_tools(golden, tmp_path):
    """`gpuga construct_train_in < input.txt` and `gpuga expand222` in the
    current directory, as the reference's tools run."""
    import subprocess
    import gpuga_amd as G
    d = golden("gaas")
    lst = unpack_dft(d, str(tmp_path))
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "gpuga")
    with open(lst) as f:
        subprocess.run([exe, "construct_train_in", "--delta-e", "0"], stdin=f, cwd=tmp_path,
                       check=True, timeout=120)
    G.construct_train_in(lst, tmp_path / "lib.in", delta_e=0.0)
    assert (tmp_path / "train.in").read_text() == (tmp_path / "lib.in").read_text()
    subprocess.run([exe, "expand222"], cwd=tmp_path, check=True, timeout=120)
    G.expand_train_in(tmp_path / "train.in", tmp_path / "lib.out")
    assert (tmp_path / "train.out").read_text() == (tmp_path / "lib.out").read_text()
