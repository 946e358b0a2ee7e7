"""Hand-computed checks of the oracle's checkerboard-update pieces (CheckerboardUpdateFrame,
MIStereo/AdaptMIPMFrame.cu:584-871; PlaneHypothesisRefinement_Frame :400-471;
GeneratePerturbedNormalLF MIStereo/CudaUtil.cu:105-142; TransformPDFToCDF CudaUtil.cu:491-507).

Every expected value below is written from the reference's text for an input chosen so that the
answer is known by hand (a closed form in the reference's constants, a position read off the
slot rules, an independent rotation-matrix product).  Each check takes the oracle library handle
so that the same checks can be run against deliberately mutated builds of the oracle
(tests/test_oracle_update_pins.py), which they must reject.
"""
import math

import numpy as np

import oracle as O

FLT_EPS = 2.0 ** -23


def _close(a, b, tol=1e-12):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.allclose(a, b, rtol=tol, atol=tol), (a, b)


def _slots_all(v):
    return np.full(8, v, np.float64)


def _ca(cols):
    """8 x cnt cost matrix from per-view columns of 8 slot costs."""
    return np.stack([np.asarray(c, np.float64) for c in cols], axis=1)


def check_probabilities_iter0(L):
    """Threshold 0.8 at iter 0; count > 2 and count_false < 3 at their boundaries; the
    exp(-c^2/0.18) average, the exp(-thr^2/0.32) fallback and 0."""
    flag = np.ones(8, bool)
    cols = [
        [0.3] * 3 + [1.0] * 5,                 # count 3 -> exp(-0.09/0.18)
        [0.3] * 2 + [1.0] * 6,                 # count 2 -> fallback exp(-0.64/0.32)
        [0.3] * 5 + [1.5] * 3,                 # count_false 3 -> 0
        [0.8] * 8,                             # c == thr is not below it -> fallback
        [0.1, 0.2, 0.3, 0.4, 1.2, 1.2, 1.3, 1.3],  # 1.2 is not > 1.2: count_false 2
        [0.3] * 3 + [1.25] * 3 + [1.0] * 2,    # count 3 but count_false 3 -> 0
    ]
    prior = np.ones(len(cols))
    prob, cdf, vw, tsel = O.view_select(_ca(cols), flag, prior, 0,
                                        O.words_from_uniforms({}), L)
    e = [math.exp(-0.5), math.exp(-2.0), 0.0, math.exp(-2.0),
         sum(math.exp(-c * c / 0.18) for c in (0.1, 0.2, 0.3, 0.4)) / 4.0, 0.0]
    _close(prob, e)
    _close(cdf, np.cumsum(e) / sum(e))


def check_probabilities_slots_absent(L):
    """Only the slots present count: three present slots at 0.3, five absent at 0."""
    flag = np.array([1, 1, 1, 0, 0, 0, 0, 0], bool)
    cols = [[0.3, 0.3, 0.3, 0, 0, 0, 0, 0], [0.3, 0.3, 1.3, 1.3, 1.3, 1.3, 1.3, 1.3]]
    prob, _, _, _ = O.view_select(_ca(cols), flag, np.ones(2), 0, O.words_from_uniforms({}), L)
    # view 1: count 2 (not > 2), count_false 1 of the present slots -> fallback
    _close(prob, [math.exp(-0.5), math.exp(-2.0)])


def check_threshold_decay(L):
    """thr = 0.8 exp(-iter^2/90): at iter 3, 0.72 < thr < 0.73; at iter 6, 0.53 < thr < 0.54."""
    flag = np.ones(8, bool)
    for it, below, above in ((3, 0.72, 0.73), (6, 0.53, 0.54)):
        thr = 0.8 * math.exp(-it * it / 90.0)
        assert below < thr < above
        cols = [[below] * 3 + [1.0] * 5, [above] * 3 + [1.0] * 5]
        prob, _, _, _ = O.view_select(_ca(cols), flag, np.ones(2), it,
                                      O.words_from_uniforms({}), L)
        _close(prob, [math.exp(-below * below / 0.18), math.exp(-thr * thr / 0.32)])


def check_prior_weighting(L):
    """Probabilities are the cost term times the prior."""
    flag = np.ones(8, bool)
    cols = [[0.3] * 3 + [1.0] * 5] * 3
    prior = np.array([2.0, 1.2, 0.4])
    prob, _, _, _ = O.view_select(_ca(cols), flag, prior, 0, O.words_from_uniforms({}), L)
    _close(prob, math.exp(-0.5) * prior)


def check_priors(L):
    """+0.9 for a set bit and +0.1 for a clear one over the direct neighbours that exist."""
    W = 16
    sel = np.zeros((W, W), np.uint32)
    sel[4, 5] = 0b101          # up of (5, 5)
    sel[6, 5] = 0b001          # down
    sel[5, 4] = 0              # left
    sel[5, 6] = 0b010 | 1 << 7  # right (bit 7 is beyond cnt)
    _close(O.view_priors(W, 5, 5, sel, 3, L), [2.0, 1.2, 1.2])
    # left border: no left neighbour
    sel[:] = 0
    sel[4, 0] = 0b11
    sel[6, 0] = 0b01
    sel[5, 1] = 0b10
    _close(O.view_priors(W, 0, 5, sel, 2, L), [0.9 + 0.9 + 0.1, 0.9 + 0.1 + 0.9])
    # corner (0, 0): right (1, 0) and down (0, 1) only
    sel[:] = 0
    sel[0, 1] = 0b1
    sel[1, 0] = 0b0
    _close(O.view_priors(W, 0, 0, sel, 2, L), [1.0, 0.2])
    # second row: the up neighbour (5, 0) exists
    sel[:] = 0
    sel[0, 5] = 0b01
    _close(O.view_priors(W, 5, 1, sel, 2, L), [0.9 + 0.3, 0.4])
    # bottom-right corner: up and left
    sel[:] = 0
    sel[W - 2, W - 1] = 0b11
    sel[W - 1, W - 2] = 0b11
    _close(O.view_priors(W, W - 1, W - 1, sel, 2, L), [1.8, 1.8])


def _k_word(k):
    """Word whose uniform is (k + 0.5) 2^-24."""
    return np.uint32(int(k) << 8)


def check_cdf_draws(L):
    """CDF (0.25, 0.5, 1) from the fallback term and priors (1, 1, 2); each draw r = u - FLT_EPSILON
    picks the first view with CDF > r (strict)."""
    flag = np.ones(8, bool)
    cols = [[1.0] * 8] * 3      # count 0, count_false 0 -> fallback for every view
    words = O.words_from_uniforms({})
    ks = [2 ** 22,           # u = 0.25 + 2^-25: r < 0.25 only with the epsilon -> view 0
          2 ** 22 + 1,       # r = 0.25 - 2^-25 -> view 0
          2 ** 22 + 2,       # r = 0.25 + 2^-25 -> view 1
          0,                 # u = 2^-25, r < 0 -> view 0
          2 ** 23 - 1,       # r just below 0.5 -> view 1
          2 ** 23 + 2,       # r = 0.5 + 2^-25 -> view 2
          2 ** 24 - 1] + [2 ** 23 + 2 ** 21] * 8   # u ~ 0.625 -> view 2
    for s, k in enumerate(ks):
        words[s] = _k_word(k)
    prob, cdf, vw, tsel = O.view_select(_ca(cols), flag, np.array([1.0, 1.0, 2.0]), 0, words, L)
    _close(prob, math.exp(-2.0) * np.array([1.0, 1.0, 2.0]))
    _close(cdf, [0.25, 0.5, 1.0])
    _close(vw, [3.0, 2.0, 10.0])
    assert tsel == 0b111, tsel
    # every draw in one view: only that bit is set and it carries all 15
    for s in range(15):
        words[s] = _k_word(2 ** 23 + 2 ** 21)
    _, _, vw, tsel = O.view_select(_ca(cols), flag, np.array([1.0, 1.0, 2.0]), 0, words, L)
    _close(vw, [0.0, 0.0, 15.0])
    assert tsel == 0b100, tsel


def check_all_views_fallback(L):
    """No probability mass (every view has >= 3 costs above 1.2): all views, weight 1."""
    flag = np.ones(8, bool)
    cols = [[1.3] * 8] * 4
    prob, _, vw, tsel = O.view_select(_ca(cols), flag, np.ones(4), 2, O.words_from_uniforms({}), L)
    _close(prob, np.zeros(4))
    _close(vw, np.ones(4))
    assert tsel == 0b1111, tsel


def _rot(a1, a2, a3):
    c, s = math.cos, math.sin
    Rx = np.array([[1, 0, 0], [0, c(a1), -s(a1)], [0, s(a1), c(a1)]])
    Ry = np.array([[c(a2), 0, s(a2)], [0, 1, 0], [-s(a2), 0, c(a2)]])
    Rz = np.array([[c(a3), -s(a3), 0], [s(a3), c(a3), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _expected_perturbed(h, px, py, pert, words):
    a = [(O.uniform_of_word(words[20 + i]) - 0.5) * pert for i in range(3)]
    n = np.asarray(h[:3], np.float64)
    o = _rot(*a) @ n
    view = np.array([px, py, 1.0]) / math.sqrt(px * px + py * py + 1.0)
    if o @ view >= 0:
        o = n.copy()
    o = o / np.linalg.norm(o)
    return np.array([o[0], o[1], o[2], h[3]])


def check_perturbed_normal(L):
    """R = Rz(a3) Ry(a2) Rx(a1) with a_i = (u_i - 0.5) * perturbation, reverted to the input
    normal when the rotated one faces away from the view ray (n . view >= 0)."""
    rng = np.random.default_rng(11)
    reverted = kept = 0
    for t in range(200):
        n = rng.normal(size=3)
        n[2] = -abs(n[2]) - 0.2
        n /= np.linalg.norm(n)
        h = np.array([n[0], n[1], n[2], 5.0 + t * 0.01])
        px, py = int(rng.integers(0, 16)), int(rng.integers(0, 16))
        pert = [0.02 * math.pi, 0.6, 2.5][t % 3]
        words = O.words_from_uniforms({20: rng.uniform(), 21: rng.uniform(), 22: rng.uniform()})
        got = O.perturbed_normal(h, px, py, pert, words, L)
        exp = _expected_perturbed(h, px, py, pert, words)
        _close(got, exp, 1e-12)
        if np.allclose(exp[:3], n):
            reverted += 1
        else:
            kept += 1
    assert reverted > 5 and kept > 100, (reverted, kept)


def _marsaglia(words, base):
    """GenerateRandomNormal (Marsaglia 1972) on the word pairs base+2a, base+2a+1."""
    for a in range(16):
        ta = 2 * (int(words[base + 2 * a]) >> 8) + 1 - 2 ** 24
        tb = 2 * (int(words[base + 2 * a + 1]) >> 8) + 1 - 2 ** 24
        if ta * ta + tb * tb < 2 ** 48:
            q1, q2 = ta * 2.0 ** -24, tb * 2.0 ** -24
            s = q1 * q1 + q2 * q2
            n = np.array([2 * q1 * math.sqrt(1 - s), 2 * q2 * math.sqrt(1 - s), 1 - 2 * s])
            return n / np.linalg.norm(n)
    raise AssertionError("no accepted pair")


def check_refine_candidates(L, pm):
    """(n, d_rand), (n_rand, d), (n_rand, d_rand), (n_pert, d), (n, d_pert) with d_rand local
    (+-1 around d) below cost 0.1 and uniform over [depth_min, depth_max] from 0.1 on, d_pert in
    [0.98 d, 1.02 d] and n_pert at perturbation 0.02 pi."""
    n = np.array([0.2, -0.1, -0.9])
    n /= np.linalg.norm(n)
    best = np.array([n[0], n[1], n[2], 6.0])
    words = O.words_from_uniforms({16: 0.75, 17: 0.25, 20: 0.9, 21: 0.1, 22: 0.7})
    # Marsaglia: first pair outside the unit disc (rejected), second accepted
    words[24] = _k_word(2 ** 24 - 1)
    words[25] = _k_word(2 ** 24 - 1)
    words[26] = _k_word(2 ** 22)
    words[27] = _k_word(3 * 2 ** 21)
    nr = _marsaglia(words, 24)
    u16, u17 = O.uniform_of_word(words[16]), O.uniform_of_word(words[17])
    npert = _expected_perturbed(best, 4, 9, 0.02 * math.pi, words)
    for bcost, local in ((0.05, True), (0.0999999, True), (0.1, False), (0.7, False)):
        got = O.refine_candidates(pm, 4, 9, best, bcost, words, L)
        dr = (u16 - 0.5) * 2.0 + 6.0 if local else u16 * (pm.depth_max - pm.depth_min) + pm.depth_min
        dp = u17 * (1.02 * 6.0 - 0.98 * 6.0) + 0.98 * 6.0
        exp = np.array([[n[0], n[1], n[2], dr], [nr[0], nr[1], nr[2], 6.0],
                        [nr[0], nr[1], nr[2], dr], [npert[0], npert[1], npert[2], 6.0],
                        [n[0], n[1], n[2], dp]])
        _close(got, exp, 1e-9)


def _pos(x, y, W=16):
    return y * W + x


def check_propagation_candidates(L):
    """Slot positions read off CheckerboardUpdateFrame's rules on a 16 x 16 MI."""
    W = 16
    cost = np.ones((W, W))
    # interior pixel, flat costs: the base positions
    fl, pos, dup = O.propagation_candidates(W, 8, 8, cost, L)
    assert fl.all()
    base = [(8, 7), (8, 5), (8, 9), (8, 11), (7, 8), (5, 8), (9, 8), (11, 8)]
    assert list(pos) == [_pos(*p) for p in base], pos
    assert (dup == -1).all()
    # minima at the alternatives: near-up diagonal (9,5), far-up step (8,4), near-left (5,9),
    # near-down diagonal (7,11), far-down step (8,12), near-right (11,7), far-right step (12,8)
    c = np.ones((W, W))
    for (x, y), v in (((9, 5), 0.1), ((8, 4), 0.2), ((5, 9), 0.05), ((7, 11), 0.3),
                      ((8, 12), 0.4), ((11, 7), 0.5), ((12, 8), 0.6), ((4, 8), 0.7)):
        c[y, x] = v
    fl, pos, dup = O.propagation_candidates(W, 8, 8, c, L)
    exp = [(9, 5), (8, 4), (7, 11), (8, 12), (5, 9), (4, 8), (11, 7), (12, 8)]
    assert list(pos) == [_pos(*p) for p in exp], [divmod(int(q), W)[::-1] for q in pos]
    # the other near-up diagonal (un - 2W - 1)
    c = np.ones((W, W))
    c[5, 7] = 0.1
    fl, pos, _ = O.propagation_candidates(W, 8, 8, c, L)
    assert pos[0] == _pos(7, 5), pos[0]
    # equal minima: the first candidate in the reference's order is kept
    c = np.ones((W, W))
    c[6, 8] = 0.2
    c[5, 7] = 0.2
    c[5, 9] = 0.2
    fl, pos, _ = O.propagation_candidates(W, 8, 8, c, L)
    assert pos[0] == _pos(8, 6), pos[0]
    # near the top border: y = 1 copies the near-up slot into the far-up slot
    fl, pos, dup = O.propagation_candidates(W, 3, 1, cost, L)
    assert fl.all() and dup[1] == 0 and pos[1] == pos[0] == _pos(3, 0)
    assert pos[5] == _pos(0, 1) and dup[5] == -1
    fl, pos, dup = O.propagation_candidates(W, 8, 2, cost, L)
    assert fl.all() and dup[1] == 0 and pos[1] == pos[0] == _pos(8, 1)
    fl, pos, dup = O.propagation_candidates(W, 8, 3, cost, L)
    assert dup[1] == -1 and pos[1] == _pos(8, 0)
    c = np.ones((W, W))
    c[2, 0] = 0.1
    fl, pos, dup = O.propagation_candidates(W, 3, 1, c, L)
    assert pos[4] == _pos(0, 2), pos[4]     # near-left alternative ln - 2 + W
    # corner (0, 0): up and left slots absent
    fl, pos, dup = O.propagation_candidates(W, 0, 0, cost, L)
    assert list(fl) == [False, False, True, True, False, False, True, True]
    assert pos[3] == _pos(0, 3) and pos[7] == _pos(3, 0)
    # near the bottom-right border: far-down and far-right copy their near slots
    fl, pos, dup = O.propagation_candidates(W, 14, 14, cost, L)
    assert fl.all() and dup[3] == 2 and dup[7] == 6
    assert pos[3] == pos[2] == _pos(14, 15) and pos[7] == pos[6] == _pos(15, 14)
    # y = 3: far-up exists without its extra step (y > 3 fails)
    c = np.ones((W, W))
    c[0, 8] = 0.0
    fl, pos, dup = O.propagation_candidates(W, 8, 3, c, L)
    assert pos[1] == _pos(8, 0) and dup[1] == -1
    fl, pos, dup = O.propagation_candidates(W, 8, 4, c, L)
    assert pos[1] == _pos(8, 0), pos[1]      # (8, 1) base, (8, 0) extra step with lower cost
    c[0, 8] = 1.0
    fl, pos, dup = O.propagation_candidates(W, 8, 4, c, L)
    assert pos[1] == _pos(8, 1)


def all_checks(pm):
    return [
        ("probabilities_iter0", check_probabilities_iter0),
        ("probabilities_slots_absent", check_probabilities_slots_absent),
        ("threshold_decay", check_threshold_decay),
        ("prior_weighting", check_prior_weighting),
        ("priors", check_priors),
        ("cdf_draws", check_cdf_draws),
        ("all_views_fallback", check_all_views_fallback),
        ("perturbed_normal", check_perturbed_normal),
        ("refine_candidates", lambda L: check_refine_candidates(L, pm)),
        ("propagation_candidates", check_propagation_candidates),
    ]
