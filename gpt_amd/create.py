"""g.create (lib/gpt/create/__init__.py:25-28)."""


def point(src, pos):
    """zero the field, then the unit matrix at site pos = [x, y, z, t]"""
    src[:] = 0
    src[tuple(pos)] = src.identity()
    return src
