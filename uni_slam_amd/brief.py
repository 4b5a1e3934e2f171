"""The rotated-BRIEF test pattern as a (256, 4) int array (ax, ay, bx, by),
read from csrc/brief_pattern.h, the table the kernels are compiled with."""
import os
import re

import numpy as np

_HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "brief_pattern.h")


def brief_pattern() -> np.ndarray:
    text = open(_HEADER).read()
    body = text[text.index("kOrbfeBriefPattern"):]
    body = body[body.index("{") + 1:body.rindex("};")]
    vals = [int(t) for t in re.findall(r"-?\d+", body)]
    assert len(vals) == 1024
    return np.array(vals, np.int64).reshape(256, 4)
