"""Selects the log table of libm.h (written to log_table.txt).

The double log of the GNU C library 2.35 splits [0x1.6p-1, 0x1.6p0) into
128 subintervals; subinterval i uses 1/c (a double near 1/center) and
log(c) rounded to a multiple of 2^-43, with c chosen so that log(c) lies
within 2^-66 of that multiple. Here:

1. every double 1/c within 2^29 ulps of 1/center is scanned (fixed point
   forward differences of 2^43 log(c), 128-bit) for the 2^-66 condition;
2. the survivors (about 250 per subinterval) are re-evaluated with 256-bit
   logarithms;
3. each survivor is compared with the host C library's log() on 3e5
   random arguments of its subinterval (x = 2^k z, |k| <= 20); the entry
   kept is the one candidate that agrees on every argument.

Needs gcc, mpmath and about a minute on 8 cores.
"""
import collections
import os
import struct
import subprocess
import tempfile

import mpmath as mp

mp.mp.prec = 256
HERE = os.path.dirname(os.path.abspath(__file__))
OFF = 0x3fe6000000000000
J = 1 << 29

SCAN_C = r"""
#include <stdio.h>
#include <stdlib.h>
typedef unsigned __int128 u128;
static u128 parse(const char *s){ u128 v=0; while(*s>='0'&&*s<='9'){ v=v*10+(*s-'0'); s++; } return v; }
int main(int argc, char **argv){
  FILE *f = fopen(argv[1], "r"); char line[4096]; int which = atoi(argv[2]), nth = atoi(argv[3]);
  const u128 TH = (u128)1 << 105;
  while (fgets(line, sizeof line, f)) {
    int i; unsigned long long inv0; char a[4][64];
    sscanf(line, "%d %llu %s %s %s %s", &i, &inv0, a[0], a[1], a[2], a[3]);
    if (i % nth != which) continue;
    u128 F = parse(a[0]), D1 = parse(a[1]), D2 = parse(a[2]), D3 = parse(a[3]);
    for (long j = -(1L << 29); j <= (1L << 29); j++) {
      if (F < TH || F > (u128)0 - TH) printf("%d %llu %ld\n", i, inv0, j);
      F += D1; D1 += D2; D2 += D3;
    }
  }
  return 0;
}
"""

VERIFY_C = r"""
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include <math.h>
static inline uint64_t asu(double f){uint64_t u; memcpy(&u,&f,8); return u;}
static inline double asd(uint64_t u){double f; memcpy(&f,&u,8); return f;}
static const double A[5] = {-0x1.0000000000001p-1, 0x1.555555551305bp-2, -0x1.fffffffeb459p-3,
                            0x1.999b324f10111p-3, -0x1.55575e506c89fp-3};
static const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
static double lg(double x, double invc, double logc){
  uint64_t ix = asu(x), tmp = ix - 0x3fe6000000000000ULL;
  int k = (int64_t)tmp >> 52;
  double z = asd(ix - (tmp & 0xfffULL << 52)), r = fma(z, invc, -1.0), kd = k;
  double w = fma(kd, Ln2hi, logc), hi = w + r, lo = fma(kd, Ln2lo, w - hi + r), r2 = r * r;
  double p = fma(r2, fma(r, A[4], A[3]), fma(r, A[2], A[1]));
  return fma(r * r2, p, fma(r2, A[0], lo)) + hi;
}
static uint64_t s = 88172645463325252ull;
static inline uint64_t rnd(void){ s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
int main(int argc, char **argv){
  FILE *f = fopen(argv[1], "r"); int i; unsigned long long a, b;
  while (fscanf(f, "%d %llu %llu", &i, &a, &b) == 3) {
    int bad = 0;
    for (int t = 0; t < 300000 && !bad; t++) {
      uint64_t z = 0x3fe6000000000000ULL + ((uint64_t)i << 45) + (rnd() & ((1ULL << 45) - 1));
      int k = (int)(rnd() % 41) - 20; if (k == 0) k = 3;
      double x = asd(z + ((uint64_t)(int64_t)k << 52));
      if (lg(x, asd(a), asd(b)) != log(x)) bad = 1;
    }
    if (!bad) printf("%d %llu %llu\n", i, a, b);
  }
  return 0;
}
"""


def asd(u):
    return struct.unpack('<d', struct.pack('<Q', u))[0]


def asu(d):
    return struct.unpack('<Q', struct.pack('<d', d))[0]


def seeds(path):
    m = 1 << 128
    with open(path, "w") as f:
        for i in range(128):
            zlo, zhi = asd(OFF + (i << 45)), asd(OFF + ((i + 1) << 45))
            inv0 = float(2 / (mp.mpf(zlo) + mp.mpf(zhi)))
            u = mp.mpf(2) ** (((asu(inv0) >> 52) - 1023) - 52)
            vals = [-mp.log(mp.mpf(inv0) + (k - J) * u) * mp.mpf(2) ** 43
                    for k in range(4)]
            d = [vals]
            for _ in range(3):
                d.append([d[-1][q + 1] - d[-1][q] for q in range(len(d[-1]) - 1)])
            fx = [int(mp.floor((x[0] - mp.floor(x[0])) * m)) % m for x in d]
            f.write("%d %d %s\n" % (i, asu(inv0), " ".join(map(str, fx))))


def main(threads=8):
    tmp = tempfile.mkdtemp()
    for name, src in (("scan", SCAN_C), ("verify", VERIFY_C)):
        with open(os.path.join(tmp, name + ".c"), "w") as f:
            f.write(src)
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o",
                               os.path.join(tmp, name),
                               os.path.join(tmp, name + ".c"), "-lm"])
    seeds(os.path.join(tmp, "seeds.txt"))
    procs = [subprocess.Popen([os.path.join(tmp, "scan"),
                               os.path.join(tmp, "seeds.txt"), str(t),
                               str(threads)], stdout=subprocess.PIPE, text=True)
             for t in range(threads)]
    cands = collections.defaultdict(list)
    for p in procs:
        out, _ = p.communicate()
        for line in out.split("\n"):
            if not line:
                continue
            i, inv0, j = map(int, line.split())
            invc = asd(inv0 + j)
            lg = -mp.log(mp.mpf(invc))
            logc = mp.nint(lg * mp.mpf(2) ** 43) / mp.mpf(2) ** 43
            if abs(lg - logc) < mp.mpf(2) ** -66 and float(lg) == float(logc):
                cands[i].append((asu(invc), asu(float(logc))))
    with open(os.path.join(tmp, "cands.txt"), "w") as f:
        for i in range(128):
            for a, b in cands[i]:
                f.write("%d %d %d\n" % (i, a, b))
    out = subprocess.check_output([os.path.join(tmp, "verify"),
                                   os.path.join(tmp, "cands.txt")], text=True)
    rows = [tuple(map(int, l.split())) for l in out.split("\n") if l]
    assert [r[0] for r in rows] == list(range(128)), "not one entry per subinterval"
    with open(os.path.join(HERE, "log_table.txt"), "w") as f:
        for r in rows:
            f.write("%d %d %d\n" % r)


if __name__ == "__main__":
    main()
