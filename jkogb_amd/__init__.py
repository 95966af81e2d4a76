"""jkogb_amd -- the J-Kogbetliantz hyperbolic SVD (venovako/JKogb's dynamic
ordering step) on one AMD MI355X: hand-written gfx950 kernels behind the C ABI
of include/jkg.h (libjkg.so), with this package as its Python front-end.

    from jkogb_amd import jkg
    with jkg.Context() as ctx:
        out = ctx.dsolve(Y, J, n2)      # YU, ZZ, SY as djk.exe writes them
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libjkg.so")

__all__ = ["HERE", "LIB"]
