"""Writes wigner_d_large.json: Wigner d^l_mn(beta) at the sampled beta of the
full-size transforms (L = 256 MW and MW_SS, L = 1024 MW), for the sparse
coefficient sets the full-size GPU tests synthesise.

Each value is Wigner's explicit sum
  d^l_{mn}(b) = sum_k (-1)^{k-n+m} sqrt((l+m)!(l-m)!(l+n)!(l-n)!)
                / ((l+n-k)! k! (l-k-m)! (k-n+m)!) cos(b/2)^{2l-2k+n-m} sin(b/2)^{2k-n+m}
evaluated in mpmath at 2l*log10(2) + 40 digits (the terms cancel by up to
2^{2l}), with b = (2j+1) pi / (2L-1) (MW) or pi j / L (MW_SS) computed in
the same precision, then rounded once to a double.  The sum and the grid
angles are textbook definitions, independent of any recursion: these values
pin the Risbo and Trapani tables and the whole GPU synthesis at full size.

Run: python tests/golden/make_wigner_fixtures.py   (about 10 s on 8 cores)
"""
import json
import math
import os
import sys
from multiprocessing import Pool

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "wigner_d_large.json")

# sparse coefficient sets (l, m, n) of the full-size tests
COEFS_256 = sorted({(255, 255, -255), (255, -3, 254), (200, -17, 33), (128, 0, 0), (97, 96, -50),
                    (1, 1, -1), (254, -200, 0), (60, 5, 60),
                    # real-signal set (n >= 0)
                    (255, 255, 255), (97, 96, 50), (1, 1, 1), (254, -200, 7)})
# L = 1024: the non-zero h_lmn of the S^2 -> SO(3) convolution test (N = 8)
# and the coefficients of the standalone N = 4 transform
FLM_1024 = [(1023, 1023), (1023, -500), (777, 3), (512, -512), (100, 17), (7, -2)]
GLM_1024_N8 = [(1023, 7), (1023, -3), (777, 0), (512, 5), (100, -7), (7, 2), (7, -7)]
COEFS_1024_N4 = [(1023, -1023, 3), (1023, 2, -3), (900, -450, 0), (513, 511, 1), (64, -64, -2), (3, 3, 3)]


def conv_coefs():
    out = []
    for (l, m) in FLM_1024:
        for (lg, n) in GLM_1024_N8:
            if lg == l:
                out.append((l, m, n))
    return out


COEFS_1024 = sorted(set(conv_coefs()) | set(COEFS_1024_N4))


def beta_list(L, ss, count=20):
    """count sample indices: both ends and an even spread between them"""
    nb = L + 1 if ss else L
    js = {0, nb - 1}
    k = 1
    while len(js) < count:
        js.add((k * 7919) % nb)
        k += 1
    return sorted(js)


def d_value(args):
    l, m, n, L, ss, j = args
    dps = int(2 * l * math.log10(2)) + 40
    with mp.workdps(dps):
        beta = mp.pi * j / L if ss else (2 * j + 1) * mp.pi / (2 * L - 1)
        c = mp.cos(beta / 2)
        s = mp.sin(beta / 2)
        tot = mp.mpf(0)
        for k in range(0, 2 * l + 1):
            a, b, cc, e = l + n - k, k, l - k - m, k - n + m
            if min(a, b, cc, e) < 0:
                continue
            den = math.factorial(a) * math.factorial(b) * math.factorial(cc) * math.factorial(e)
            t = c ** (2 * l - 2 * k + n - m) * s ** (2 * k - n + m) / den
            tot += -t if (k - n + m) % 2 else t
        rad = math.factorial(l + m) * math.factorial(l - m) * math.factorial(l + n) * math.factorial(l - n)
        return float(tot * mp.sqrt(rad))


def main():
    cfgs = {"mw256": (256, 0, COEFS_256), "mwss256": (256, 1, COEFS_256), "mw1024": (1024, 0, COEFS_1024)}
    out = {"_source": "tests/golden/make_wigner_fixtures.py (mpmath %s, Wigner's explicit sum)" % mp.__version__,
           "flm_1024": FLM_1024, "glm_1024_n8": GLM_1024_N8, "coefs_1024_n4": COEFS_1024_N4}
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for key, (L, ss, coefs) in cfgs.items():
            bs = beta_list(L, ss)
            jobs = [(l, m, n, L, ss, j) for (l, m, n) in coefs for j in bs]
            vals = pool.map(d_value, jobs, chunksize=1)
            table = {}
            it = iter(vals)
            for (l, m, n) in coefs:
                table["%d,%d,%d" % (l, m, n)] = [next(it) for _ in bs]
            out[key] = {"L": L, "ss": ss, "b": bs, "d": table}
            print(key, len(jobs), "values", flush=True)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    return 0


if __name__ == "__main__":
    sys.exit(main())
