"""Gradient of the summed forward log-likelihood of one profile over DNA
sequences, as a torch.autograd.Function.

F(emission logits, transition logits) = sum_s ln Z_s, where Z_s is the total
of all local alignments of sequence s (sum_{i,j} prefixSeed(j) W1(i,j) over
the reference's recurrences, dummest-dp-forward.hh:9-250 and
dummest-dp-backward.hh:444-615), and the profile is built from the logits the
way finalizeProfile builds it from HMM probabilities (dummest-profile.hh:166-224):

* emission logits (M x 20): node k+1's letter probabilities p = softmax(row k);
  the codon emission of profile row k is (1 - STOP) p_aa / #synonymous codons
  (U as C, O as K, a codon with another letter 1/64, stop codons STOP/3);
* transition logits ((M+1) x 7, HMMER order m->m m->i m->d i->m i->i d->m d->d):
  per state a softmax, giving MtoI, MtoD, ItoI, DtoD, then alpha' =
  MtoI keep (1 - ItoI), beta' = ItoI, delta' = MtoD keep (1 - DtoD[i+1]),
  eps' = DtoD (1 - DtoD[i+1]) / (1 - DtoD), enter = 1 - MtoI keep - MtoD keep -
  INSERT1 - INSERT2 - DELETE1 - DELETE2 with keep = 1 - INSERT1 - INSERT2 -
  DELETE1 - DELETE2.

The background letter probabilities and with them the null model are inputs
held fixed (a search uses the background the profile was read with).

The GPU (dm_grad) returns per row the expected uses of each Params entry and
of each codon's emission, d ln Z / d ln theta; the chain rule to the logits is
applied here in FP64.
"""
from __future__ import annotations

import numpy as np

from . import DEFAULT_RATES, WIDTH

AA = "ACDEFGHIKLMNPQRSTVWY"
# the standard genetic code, codons in ACGT order (index 16 b0 + 4 b1 + b2)
_CODE = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
CODON_AA = np.array([AA.index(c) if c != "*" else -1 for c in _CODE], np.int64)  # -1: stop
SYN = np.array([int((CODON_AA == k).sum()) for k in range(20)], np.float64)


def _softmax(x, axis=-1):
    z = np.exp(x - x.max(axis=axis, keepdims=True))
    return z / z.sum(axis=axis, keepdims=True)


def transitions(trans_logits):
    """MtoI, MtoD, ItoI, DtoD per node from (M+1) x 7 logits."""
    t = np.asarray(trans_logits, np.float64)
    pm, pi, pd = _softmax(t[:, 0:3]), _softmax(t[:, 3:5]), _softmax(t[:, 5:7])
    return pm[:, 1], pm[:, 2], pi[:, 1], pd[:, 1], (pm, pi, pd)


def profile_from_logits(emis_logits, trans_logits, bg, rates=DEFAULT_RATES):
    """(values (M+1) x 29, params (M+2) x 11) in FP64 as finalizeProfile forms
    them; bg (29) is kept as given."""
    el = np.asarray(emis_logits, np.float64)
    M = el.shape[0]
    MtoI, MtoD, ItoI, DtoD, _ = transitions(trans_logits)
    assert len(MtoI) == M + 1
    i1, i2, d1, d2, stop = rates[0], rates[1], rates[2], rates[3], rates[4]
    keep = 1 - i1 - i2 - d1 - d2
    values = np.zeros((M + 1, WIDTH))
    params = np.zeros((M + 2, 11))
    p = _softmax(el)
    for r in range(M + 1):
        alpha, beta = MtoI[r] * keep, ItoI[r]
        params[r, 0:4] = [alpha * (1 - beta), i1, i2, beta]
        values[r, 0:2] = [MtoI[r], ItoI[r]]
        if r == M:  # the end row holds the background letters (finalizeProfile's `end`)
            values[r, 4:24] = bg[4:24]
            break
        delta, eps, eps1 = MtoD[r] * keep, DtoD[r], DtoD[r + 1]
        params[r, 6:11] = [delta * (1 - eps1), d1, d2, eps * (1 - eps1) / (1 - eps),
                           1 - alpha - i1 - i2 - delta - d1 - d2]
        values[r, 2:4] = [delta, eps]
        values[r, 4:24] = (1 - stop) * p[r] / SYN
        values[r, 24], values[r, 25] = values[r, 5], values[r, 12]  # U = C, O = K
        values[r, 26], values[r, 27], values[r, 28] = 1 / 64, stop / 3, 0.0
    return values, params


def logit_grads(counts, emis_logits, trans_logits, rates=DEFAULT_RATES):
    """Chain rule from dm_grad's counts ((M+1) x 76) to (d F / d emission
    logits (M x 20), d F / d transition logits ((M+1) x 7)), and the codon-level
    emission gradient d F / d ln e (M x 64)."""
    c = np.asarray(counts, np.float64)
    el = np.asarray(emis_logits, np.float64)
    M = el.shape[0]
    MtoI, MtoD, ItoI, DtoD, (pm, pi, pd) = transitions(trans_logits)
    keep = 1 - rates[0] - rates[1] - rates[2] - rates[3]
    a0, b0, d0, eps, enter = c[:, 0], c[:, 3], c[:, 6], c[:, 9], c[:, 10]
    alpha, delta = MtoI * keep, MtoD * keep
    ent = 1 - alpha - delta - rates[0] - rates[1] - rates[2] - rates[3]
    ent_share = np.where(enter != 0, enter / np.where(ent != 0, ent, 1.0), 0.0)
    g_mi = a0 - ent_share * alpha
    g_ii = -a0 * ItoI / (1 - ItoI) + b0
    g_md = d0 - ent_share * delta
    g_dd = eps / (1 - DtoD)
    g_dd[1:] += -(d0[:-1] + eps[:-1]) * DtoD[1:] / (1 - DtoD[1:])
    gm = np.stack([np.zeros(M + 1), g_mi, g_md], 1)
    gi = np.stack([np.zeros(M + 1), g_ii], 1)
    gd = np.stack([np.zeros(M + 1), g_dd], 1)
    gt = np.concatenate([gm - pm * gm.sum(1, keepdims=True), gi - pi * gi.sum(1, keepdims=True),
                         gd - pd * gd.sum(1, keepdims=True)], 1)
    codon = c[:M, 11:11 + 64]
    n_aa = np.zeros((M, 20))
    for k in range(20):
        n_aa[:, k] = codon[:, CODON_AA == k].sum(1)
    p = _softmax(el)
    ge = n_aa - p * n_aa.sum(1, keepdims=True)
    return ge, gt, codon


class SummedLogLikelihood:
    """F(emission logits, transition logits) = sum_s ln Z_s on the GPU, with
    its gradient (torch.autograd.Function; see forward()).

    engine: a dummest_amd.Engine; seqs: DNA strings; bg: the 29 background
    values of the profile (held fixed); rates as dm_set_rates."""

    def __init__(self, engine, seqs, bg, rates=DEFAULT_RATES, weights=None):
        self.engine, self.seqs, self.bg = engine, list(seqs), np.asarray(bg, np.float64)
        self.rates, self.weights = tuple(rates), weights
        self.pid = None  # the engine's slot for the current profile, reused by every evaluation
        self.set_id = None  # the engine's registered copy of the sequences

    def evaluate(self, emis_logits, trans_logits):
        """(F, d F / d emission logits, d F / d transition logits, per-sequence ln Z)."""
        values, params = profile_from_logits(emis_logits, trans_logits, self.bg, self.rates)
        M = values.shape[0] - 1
        params = np.vstack([params, np.zeros((1, 11))]) if params.shape[0] == M + 1 else params
        if self.pid is None:
            self.pid = self.engine.add_profile(M, values, params, self.bg)
        else:
            self.engine.set_profile(self.pid, M, values, params, self.bg)
        if self.set_id is None:  # decoded and uploaded once for every evaluation
            self.set_id = self.engine.grad_set(self.seqs)
        lnz, counts = self.engine.grad_set_counts(self.pid, self.set_id, self.weights)
        ge, gt, _ = logit_grads(counts, emis_logits, trans_logits, self.rates)
        w = np.ones(len(lnz)) if self.weights is None else np.asarray(self.weights, np.float64)
        return float((w * lnz).sum()), ge, gt, lnz

    def function(self):
        """A torch.autograd.Function class bound to this engine and these sequences."""
        import torch

        outer = self

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, emis_logits, trans_logits):
                F, ge, gt, _ = outer.evaluate(emis_logits.detach().cpu().double().numpy(),
                                              trans_logits.detach().cpu().double().numpy())
                ctx.save_for_backward(torch.from_numpy(ge), torch.from_numpy(gt))
                ctx.dtypes = (emis_logits.dtype, trans_logits.dtype, emis_logits.device)
                return torch.tensor(F, dtype=torch.float64, device=emis_logits.device)

            @staticmethod
            def backward(ctx, grad_out):
                ge, gt = ctx.saved_tensors
                de, dt, dev = ctx.dtypes
                g = grad_out.detach().cpu().double()
                return (g * ge).to(dtype=de, device=dev), (g * gt).to(dtype=dt, device=dev)

        return _F

