"""Drop-in replacement of the reference extension module `isinglib`
(mcsolver/mcMain.py imports it by this name).  Runs on the MI355X."""
from .dropin import ising_main


def MCMainFunction(*args):
    """isinglib.MCMainFunction (isingLib.c:259-451): 16 positional arguments
    (algorithm, initSpin, nthermal, nsweep, ninterval, maxNLinking, nlink,
    linkStrength, linkedOrb, corrOrbPair, h, rOrb, rOrbCluster,
    linkedOrb_rnorm, spinFrame, callback) -> 11-tuple."""
    if len(args) != 16:
        raise TypeError("MCMainFunction takes 16 positional arguments (%d given)" % len(args))
    return ising_main(*args)
