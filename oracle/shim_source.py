"""Print the C shim of INTEGRATION.md (its first ```c block), the file a
maintainer adds to the reference to call libmcl_hip.so; the Makefile's
dropin target compiles it."""
import os
import re
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def shim_source():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    sec = text.split("## C shim a maintainer adds to the reference", 1)[1]
    return re.search(r"```c\n(.*?)```", sec, re.S).group(1)


if __name__ == "__main__":
    sys.stdout.write(shim_source())
