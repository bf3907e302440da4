"""Point / IntPoint value types (reference: base/point.pyx, point.pxd)."""
import math


class Point:
    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def set(self, x, y, z):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def asarray(self):
        return (self.x, self.y, self.z)

    def norm(self):
        return self.x * self.x + self.y * self.y + self.z * self.z

    def length(self):
        return math.sqrt(self.norm())

    def dot(self, p):
        return self.x * p.x + self.y * p.y + self.z * p.z

    def __sub__(self, p):
        return Point(self.x - p.x, self.y - p.y, self.z - p.z)

    def __add__(self, p):
        return Point(self.x + p.x, self.y + p.y, self.z + p.z)

    def __mul__(self, k):
        return Point(self.x * k, self.y * k, self.z * k)

    def __eq__(self, p):
        return isinstance(p, Point) and self.asarray() == p.asarray()

    def __repr__(self):
        return "Point(%g, %g, %g)" % self.asarray()


class IntPoint:
    def __init__(self, x=0, y=0, z=0):
        self.x, self.y, self.z = int(x), int(y), int(z)

    def asarray(self):
        return (self.x, self.y, self.z)

    def __eq__(self, p):
        return isinstance(p, IntPoint) and self.asarray() == p.asarray()

    def __hash__(self):
        return hash(self.asarray())

    def __repr__(self):
        return "IntPoint(%d, %d, %d)" % self.asarray()
