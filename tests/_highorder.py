"""The reference's high-order output files, computed from 4-site window
marginals W[o, p] = sum_s x_s [((s >> o) & 15) == p].

The reference evaluates each value with a loop over all states
(src/masterEquation.c:32-143: key_fit, error_guess, higher_order_pf_approx,
higher_order_point_func; files written at :1422-1539).  Every key_fit sum it
takes covers at most 4 consecutive sites, so it is a sum of W entries; this
module restates the file formats on top of that, the same way the C host
does, so a test can compare the text with the reference's files.
"""


def key_sum(W, n, key, off, size, width=4):
    """sum_s x_s over states whose bits off..off+size-1 equal key, from
    windows of `width` sites."""
    w = min(off, n - width)
    sh = off - w
    m = (1 << size) - 1
    acc = 0.0
    for p in range(1 << width):
        if ((p >> sh) & m) == key:
            acc += W[w][p]
    return acc


def _sums(W, n, key, ks, off):
    k_left = key + (1 << ks)
    k_right = (key << 1) + 1
    k_lr = (key << 1) + 1 + (1 << (ks + 1))
    return (key_sum(W, n, k_lr, off - 1, ks + 2), key_sum(W, n, key, off, ks),
            key_sum(W, n, k_left, off, ks + 1), key_sum(W, n, k_right, off - 1, ks + 1))


def error_guess(W, n, key, ks, off):
    lr, no, left, right = _sums(W, n, key, ks, off)
    return (lr / no - (left / no) * (right / no)) * no


def closure(W, n, key, ks, off):
    _, no, left, right = _sums(W, n, key, ks, off)
    return (left * right) / no


def files(W, n):
    """Text of the error-guess, closures and point-function files."""
    ms, mask = 2, 3
    eg = "".join("Error guess for t_%d...t_%d:\t %f\n" % (i - 1, i + ms, error_guess(W, n, mask, ms, n - ms - i))
                 for i in range(1, n - ms))
    cl = "".join("Value for closure of t_%d...t_%d:\t %f\n" % (i - 1, i + ms, closure(W, n, mask, ms, n - ms - i))
                 for i in range(1, n - ms))
    ms, mask = 4, 15
    pf = "".join("t_%d...t_%d = %f\n" % (i, ms + i - 1, key_sum(W, n, mask, n - ms - i, ms))
                 for i in range(0, n - ms + 1))
    return {"out_error_guess": eg, "out_closures": cl, "out_high_order_point_fun": pf}
