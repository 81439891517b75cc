"""Surfacelet (decimated): placeholder until the decimated path lands."""


class Surfacelet:
    pass


class SurfaceletWorkspace:
    pass
