"""Parser for the reference solver's stdout (list-directed Fortran output,
run with FORT_FMT_RECL large so that no record wraps).

Produces the same (kind, values) event tuples the oracle reports:
  iter0:  (L0, sqrt(normder))                                 pqdts.f90:1914
  active: (nact, M*N)                                          pqdts.f90:2002
  pcg:    (j, sqrt(b2), eps, b1, v2)                           pqdts.f90:2084
  ls:     (i, L, alpha, df, xmin)  (xmin = 0 for stage 1)      pqdts.f90:429/474
  iter:   (oiter, iiter, L, xmin, conv, normder, mode, alpha, dlalpha)  pqdts.f90:2131
  stop1:  (dlalpha,)                                           pqdts.f90:1956
"""
import re

EV_ITER, EV_PCG, EV_LS, EV_ACTIVE, EV_STOP1, EV_ITER0, EV_LAMBDA = range(7)


def _f(s):
    return float(s)


def parse(text):
    ev = []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "iter" and t[3] == "L=" and len(t) > 5 and t[5] == "||der||=":
            ev.append((EV_ITER0, (_f(t[4]), _f(t[6]))))
        elif t[0] == "iter":
            # iter o i L= v max(abs(c)) v conv v ||der||= v mode= m alpha= a dlalpha= d ...
            ev.append((EV_ITER, (_f(t[1]), _f(t[2]), _f(t[4]), _f(t[6]), _f(t[8]), _f(t[10]),
                                 _f(t[12]), _f(t[14]), _f(t[16]))))
        elif t[0] == "positivity":
            ev.append((EV_ACTIVE, (_f(t[2]), _f(t[5]))))
        elif t[0] == "newton-pcg":
            ev.append((EV_PCG, (_f(t[1]), _f(t[3]), _f(t[5]), _f(t[7]), _f(t[9]))))
        elif t[0] == "line" and t[1] == "search":
            xmin = _f(t[11]) if len(t) > 11 else 0.0
            ev.append((EV_LS, (_f(t[2]), _f(t[4]), _f(t[6]), _f(t[8]), xmin)))
        elif t[0] == "stopping":
            ev.append((EV_STOP1, (_f(t[-1]),)))
    return ev
