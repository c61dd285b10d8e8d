"""Drop-in replacement of the reference extension module `xylib`
(mcsolver/mcMain.py imports it by this name).  Runs on the MI355X."""
from .dropin import on_main


def MCMainFunction(*args):
    """xylib.MCMainFunction (xyLib.c:413-809): 23 positional arguments ->
    29-tuple; slots 2, 5, 13, 16 are 0.0 and Q is 0 (xyLib.c:781-805)."""
    if len(args) != 23:
        raise TypeError("MCMainFunction takes 23 positional arguments (%d given)" % len(args))
    return on_main(2, *args)
