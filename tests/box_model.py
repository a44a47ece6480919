"""NumPy statement of the box smoother of the FP32 cycles (test
infrastructure): the library's partition of a level into boxes of at most
BOX nodes a side, and one sweep of the node-block Gauss-Seidel update
(General_matrix_functions.c:1160-1361) in the 8-colour order of the
restatement, run box by box: inside a box each colour sees the colours
before it as updated; a node of another box is read as it was before the
sweep.  With one box for the whole level the sweep is the restatement's
8-colour sweep (tests/test_box_model.py checks that against the oracle)."""
import numpy as np

BOX = 8


def axis_split(n):
    """segment starts of at most BOX nodes, as equal as they come (the
    library's partition)"""
    nseg = (n + BOX - 1) // BOX
    s, pos = [], 0
    for t in range(nseg):
        s.append(pos)
        pos += n // nseg + (1 if t < n % nseg else 0)
    return np.array(s + [n])


def box_ids(n):
    s = axis_split(n)
    ids = np.zeros(n, dtype=np.int64)
    for t in range(len(s) - 1):
        ids[s[t]:s[t + 1]] = t
    return ids


def sweep_model(K, BI, F, d, dims, boxes=True, half=False):
    """one sweep; K [nno][243] (row r, slot (dz+1)+3(dx+1)+9(dy+1), column),
    node n = k + noz (i + nox j).  half: the rows scaled by BI and rounded
    to FP16, the update BI F - (BI K) d (the mixed = 2 form's levels)"""
    noz, nox, noy = dims
    nno = noz * nox * noy
    K = K.reshape(nno, 3, 27, 3)
    if half:
        K = (K * BI.astype(np.float32).reshape(nno, 3, 1, 1)).astype(np.float16)
    K = K.astype(np.float64)
    d = d.reshape(nno, 3).copy()
    old = d.copy()
    F = F.reshape(nno, 3)
    BI = BI.reshape(nno, 3)
    if boxes:
        bz, bx, by = box_ids(noz), box_ids(nox), box_ids(noy)
    else:
        bz, bx, by = np.zeros(noz, int), np.zeros(nox, int), np.zeros(noy, int)
    kk, ii, jj = np.meshgrid(np.arange(noz), np.arange(nox), np.arange(noy), indexing="ij")
    for c in range(8):
        sel = ((kk & 1) == (c & 1)) & ((ii & 1) == ((c >> 1) & 1)) & ((jj & 1) == ((c >> 2) & 1))
        k, i, j = kk[sel], ii[sel], jj[sel]
        n = k + noz * (i + nox * j)
        Ad = np.zeros((n.size, 3))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    s = (dz + 1) + 3 * (dx + 1) + 9 * (dy + 1)
                    k2, i2, j2 = k + dz, i + dx, j + dy
                    ok = (k2 >= 0) & (k2 < noz) & (i2 >= 0) & (i2 < nox) & (j2 >= 0) & (j2 < noy)
                    k2c, i2c, j2c = np.clip(k2, 0, noz - 1), np.clip(i2, 0, nox - 1), np.clip(j2, 0, noy - 1)
                    m = k2c + noz * (i2c + nox * j2c)
                    same = (bz[k2c] == bz[k]) & (bx[i2c] == bx[i]) & (by[j2c] == by[j])
                    u = np.where(same[:, None], d[m], old[m])
                    Ad += np.where(ok[:, None], np.einsum("nrc,nc->nr", K[n, :, s, :], u), 0.0)
        d[n] += F[n] * BI[n] - Ad if half else (F[n] - Ad) * BI[n]
    return d.reshape(-1)
