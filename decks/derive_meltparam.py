"""Derivation of the decks' `meltparam` (front, tail, half-width, depth of the melt pool).

The workload texts give beam power, speed, absorptivity 0.35, Gaussian radius 50 um,
k = 20 W/mK, alpha = 5e-6 m^2/s and T0 = 300 K; the reference's analytic temperature
model needs the melt-pool semi-axes instead (S/TempField.C:71-76, 175-190).  They are the
extent of the T >= tL (1733 K) isosurface of the constant-property Gaussian moving source
(the EASM form of S/Utilities.C:228-249 with constant properties), evaluated here with
scipy quadrature.  Run: python decks/derive_meltparam.py
"""
import numpy as np
from scipy.integrate import quad
def T(x,y,z,P,v,eta,sig,k=20.0,alpha=5e-6,T0=300.0):
    rc=k/alpha
    C=eta*P/(np.pi*rc*np.sqrt(4*np.pi*alpha))
    f=lambda tp: C*tp**-0.5/(2*alpha*tp+sig**2)*np.exp(-((x+v*tp)**2+y**2)/(4*alpha*tp+2*sig**2)-z**2/(4*alpha*tp))
    val,_=quad(f,1e-12,2e-3,limit=400,points=[1e-7,1e-6,1e-5,1e-4])
    return T0+val
def dims(P,v,eta=0.35,sig=50e-6,tL=1733.0):
    xs=np.linspace(-600e-6,200e-6,801)
    # along the axis at the surface (z->0+)
    z0=1e-9
    Tx=np.array([T(x,0,z0,P,v,eta,sig) for x in xs])
    hot=xs[Tx>=tL]
    if hot.size==0: return None
    front, tail = hot.max(), -hot.min()
    # width: search y at x of max width
    best=(0,0)
    for x in np.linspace(-tail,front,40):
        ys=np.linspace(0,200e-6,201); Ty=np.array([T(x,y,z0,P,v,eta,sig) for y in ys]); w=ys[Ty>=tL]
        if w.size and w.max()>best[0]: best=(w.max(),x)
    hw,xw=best
    zs=np.linspace(1e-7,200e-6,201); Tz=np.array([T(xw,0,z,P,v,eta,sig) for z in zs]); d=zs[Tz>=tL].max() if (Tz>=tL).any() else 0
    return front,tail,hw,d
def _edge(f, a, b, tol=1e-7):
    """Bisection for the boundary between f(a) True and f(b) False."""
    while abs(b - a) > tol:
        m = .5 * (a + b)
        if f(m):
            a = m
        else:
            b = m
    return a
def dims_bisect(P,v,eta=0.35,sig=50e-6,tL=1733.0):
    """Same extents as dims() by bisection on the isosurface (a few hundred quadratures)."""
    z0=1e-9
    hot=lambda x,y,z: T(x,y,z,P,v,eta,sig)>=tL
    xs=np.linspace(-800e-6,200e-6,251)
    h=[x for x in xs if hot(x,0,z0)]
    if not h: return None
    front=_edge(lambda x: hot(x,0,z0), max(h), max(h)+4e-6)
    tail=-_edge(lambda x: hot(x,0,z0), min(h), min(h)-4e-6)
    best=(0,0)
    for x in np.linspace(-tail,front,40):
        if hot(x,0,z0):
            w=_edge(lambda y: hot(x,y,z0), 0.0, 400e-6)
            if w>best[0]: best=(w,x)
    hw,xw=best
    d=_edge(lambda z: hot(xw,0,z), 1e-7, 400e-6)
    return front,tail,hw,d
if __name__ == "__main__":
    for name,P,v in [("c0",200,0.8),("c1",250,1.0)]:
        print(name, ["%.1f um"%(1e6*q) for q in dims(P,v)])
