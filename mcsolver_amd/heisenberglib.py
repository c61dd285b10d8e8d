"""Drop-in replacement of the reference extension module `heisenberglib`
(mcsolver/mcMain.py imports it by this name).  Runs on the MI355X."""
from .dropin import on_main


def MCMainFunction(*args):
    """heisenberglib.MCMainFunction (heisenbergLib.c:478-885): 23 positional
    arguments -> 29-tuple."""
    if len(args) != 23:
        raise TypeError("MCMainFunction takes 23 positional arguments (%d given)" % len(args))
    return on_main(3, *args)
