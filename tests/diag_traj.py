"""Diagnostic (not collected): per-step deviation of the GPU FP64 trajectory
from the oracle."""
import sys, os
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import numpy as np
from j_presto_amd import engine, systems
from oracle import Oracle
for L, nw in [(24.8, 500), (39.1, 2000)]:
    s = systems.water_box(nw, L, 7)
    print(f"case L={L} n={s.n}", flush=True)
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    o.celprp(x)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    md.celprp()
    for k in range(1, 21):
        g, ee, ev = o.nonbond(x)
        gg, eeg, evg, _, _ = md.cmmdo()
        o.verlet_step(x, v, o.combine(g))
        md.runmd(1, update=1000)
        xg, vg = md.get_state()
        print(k, "dgrad %.3e" % np.abs(gg[0] - g[0]).max(), "dx %.3e dv %.3e" % (np.abs(xg - x).max(), np.abs(vg - v).max()), flush=True)
        md.set_state(x, v)   # resync to the oracle so each step is compared from the same state
        if k % 10 == 0:
            o.celprp(x)
            md.celprp()
