"""The closed-form operator generator (dgk_make_operators) and the
operator-file writer (dgk_write_operators), on the CPU.

* For p = 1..15 the generated x1D, D1D and Lift_mat equal the reference's
  operator_data/p{p}.dat files to the files' own rounding.
* For p = 16..19, where the reference ships no files, the generated arrays
  equal the closed forms computed independently with numpy's Legendre
  routines (tests/test_oracle_operators.py), and satisfy the identities
  D 1 = 0, D x = 1, W D + D^T W = diag(-1, 0, ..., 1), w^T l-/+ = 1.
* A written p{p}.dat reads back bit-exactly through the product's and the
  oracle's loaders, and dgk_load_operators falls back to the generator above
  p = 15 when the directory has no file.
"""
import numpy as np
import pytest

import dgk
import oracle
from test_oracle_operators import diff_matrix, ell, lgl


@pytest.mark.parametrize("p", range(1, 16))
def test_generated_equals_stored(p):
    gx, gD, gL = dgk.make_operators(p)
    sx, sD, sL = dgk.load_operators(p)
    assert np.abs(gx - sx).max() <= 2e-15
    assert np.abs(gD - sD).max() <= 5e-12 * np.abs(sD).max()
    assert np.abs(gL - sL).max() <= 5e-12 * np.abs(sL).max()


@pytest.mark.parametrize("p", range(16, 20))
def test_generated_beyond_the_files(p):
    x1D, Dm, Lm = dgk.make_operators(p)
    D = Dm.T  # D[i, l] = D1D(i, l)
    x, w, Pp = lgl(p)
    assert np.abs(x1D - x).max() <= 4e-15
    assert np.all(np.diff(x1D) > 0) and np.array_equal(x1D, -x1D[::-1])
    scale = np.abs(D).max()
    assert np.abs(D - diff_matrix(p, x, Pp)).max() <= 1e-12 * scale
    n = p + 1
    assert np.abs(D @ np.ones(n)).max() <= 1e-13 * scale
    assert np.abs(D @ x1D - 1.0).max() <= 1e-13 * scale
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    W = np.diag(w)
    assert np.abs(W @ D + D.T @ W - B).max() <= 1e-13 * scale
    lm, lp = ell(p, x, -1), ell(p, x, +1)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    along = [J, I, J, I, K, K]
    side = [lm, lp, lp, lm, lm, lp]
    lscale = np.abs(Lm).max()
    for f in range(6):
        assert np.abs(Lm[f].transpose(2, 1, 0) - side[f][along[f]]).max() <= 1e-12 * lscale, f
    assert abs(w @ Lm[3, 0, 0, :] - 1.0) <= 1e-13 * lscale  # w^T l- = 1
    assert abs(w @ Lm[1, 0, 0, :] - 1.0) <= 1e-13 * lscale  # w^T l+ = 1


@pytest.mark.parametrize("p", [1, 4, 15, 17])
def test_written_file_reads_back(tmp_path, p):
    dgk.write_operators(tmp_path, p)
    text = (tmp_path / f"p{p}.dat").read_text().splitlines()
    assert len(text) == 3 and all(len(t) == 25 * n for t, n in
                                  zip(text, ((p + 1), (p + 1) ** 2, 6 * (p + 1) ** 3)))
    g = dgk.make_operators(p)
    for a, b, c in zip(g, dgk.load_operators(p, tmp_path), oracle.load_operators(p, tmp_path)):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_loader_generates_above_the_files(tmp_path):
    for a, b in zip(dgk.load_operators(18, tmp_path), dgk.make_operators(18)):
        assert np.array_equal(a, b)
    with pytest.raises(dgk.DgkError) as ei:  # the reference's orders still need their files
        dgk.load_operators(7, tmp_path)
    assert ei.value.code == dgk.ERR_IO
    with pytest.raises(dgk.DgkError) as ei:
        dgk.load_operators(20)
    assert ei.value.code == dgk.ERR_UNSUPPORTED
