"""Real-valued lane-packing criteria at a point, in mpmath — TEST INFRASTRUCTURE.

The interval criteria of the reference bound real-valued area inequalities over
a cell.  This module states those inequalities at one point (la, r1, r2, r3)
from the geometry the reference's formulas describe, with independent closed
forms where the reference uses a construction:

* collision angle of two touching disks (collision_angles.cu:37-117 builds it
  from the altitude foot and two arcsines): the angle at the origin of the
  triangle with sides ca, cb (centre distances) and ab = ra + rb, by the law of
  cosines, theta = acos((ca^2 + cb^2 - ab^2) / (2 ca cb)); disks that pass each
  other radially (|ca - cb| >= ab) have angle 0;
* centre distances: inner disk la + r, outer disk 1 - r (collision_angles.cu:
  119-231 pairs them as inner_inner, outer_inner, inner_outer, outer_outer);
* semicircle angles asin(r / (1 - r)) (outer) and asin(r / (la + r)) (inner)
  (collision_angles.cu:233-246);
* disk areas pi r^2, the lane's area per radian (1 - la^2) / 2 and the
  critical ratio c = 0.5642 (values.hpp:28);
* reusable areas: the reference clamps reusable angles and the ring factor rho
  at 0 where it says so (start_configuration.cu:176,181,244-254;
  end_configuration.cu:79,95,172,201); an area that cannot be reused is 0, so
  the point form clamps every reusable area at 0 (the interval bound is then
  still a lower bound of it).

Each family's criterion is a list of checks (margin = disk area - c * used
area; the check holds when the margin is >= 0) selected by case gates (margin
of the gate's comparison).  `evaluate` returns whether the family's criterion
holds at the point and the smallest |margin| of every decision taken, scaled
by the lane area, so tests can exclude points within rounding distance of a
decision boundary."""
import mpmath as mp

mp.mp.dps = 40
C = mp.mpf(0.5642)  # the double the reference multiplies with (values.hpp:28)
PI = mp.pi
START, START_END_2D, START_END_3D, MIDDLE, END_2D, END_3D = range(6)


def theta(ca, cb, ab):
    c = (ca * ca + cb * cb - ab * ab) / (2 * ca * cb)
    if c >= 1:
        return mp.mpf(0)
    if c <= -1:
        return PI
    return mp.acos(c)


def inner_inner(la, ra, rb):
    return theta(la + ra, la + rb, ra + rb)


def outer_inner(la, ra, rb):
    return theta(1 - ra, la + rb, ra + rb)


def inner_outer(la, ra, rb):
    return theta(la + ra, 1 - rb, ra + rb)


def outer_outer(ra, rb):
    return theta(1 - ra, 1 - rb, ra + rb)


def s_out(r):
    return mp.asin(r / (1 - r))


def s_in(la, r):
    return mp.asin(r / (la + r))


def area(r):
    return PI * r * r


def pos(x):
    return x if x > 0 else mp.mpf(0)


class Decisions:
    """Collects checks and gates; `holds` is the conjunction of the checks and
    `pairs` the two sides (disk area, c x used area) of every check, in the
    reference's evaluation order."""

    def __init__(self, scale):
        self.scale = scale
        self.margins = []
        self.pairs = []
        self.holds = True

    def gate(self, margin):
        self.margins.append(abs(margin))
        return margin

    def check(self, disk, used):
        bound = C * used
        m = disk - bound
        self.pairs.append((disk, bound))
        self.margins.append(abs(m) / self.scale)
        if m < 0:
            self.holds = False
        return m >= 0


def evaluate(f, la, r1, r2, r3=None, physical=True):
    """(holds, closest decision margin, check pairs) of family f's criterion at
    the point.  physical=True clamps every reusable area at 0 (the quantity the
    interval bounds are lower bounds of); physical=False keeps the end
    families' reusable angle signed, as their formulas are written
    (end_configuration.cu:79-80, 95-96, 170-172, 199-201), for comparing values
    at point cells."""
    la, r1, r2 = mp.mpf(la), mp.mpf(r1), mp.mpf(r2)
    r3 = mp.mpf(r3) if r3 is not None else None
    ratio = (1 - la * la) / 2  # lane area per radian
    reuse_angle = pos if physical else (lambda x: x)
    d = Decisions(ratio)
    if f == START:  # start_configuration.cu:93-112, 137-162
        a12, a13, a23 = outer_inner(la, r1, r2), outer_outer(r1, r3), inner_outer(la, r2, r3)
        if d.gate(a12 + a23 - a13) > 0:  # r3 stopped by r2
            d.check(area(r1) + area(r2) / 2, ratio * (s_out(r1) + a12))
        else:  # r3 stopped by r1
            d.check(area(r1) + area(r2) + area(r3) / 2, ratio * (s_out(r1) + a13))
    elif f == START_END_2D:  # start_configuration.cu:169-188
        a1, a2, a12 = s_out(r1), s_in(la, r2), outer_inner(la, r1, r2)
        reuse = pos(a12 + a2 - a1) * pos(1 - (la + 2 * r2) ** 2) / 2
        d.check(area(r1) + area(r2), ratio * (a1 + a2 + a12) - reuse)
    elif f == START_END_3D:  # start_configuration.cu:227-265 (a13 of r1 and r2 as at :232)
        a1, a2, a3 = s_out(r1), s_in(la, r2), s_out(r3)
        a12, a13, a23 = outer_inner(la, r1, r2), outer_outer(r1, r2), inner_outer(la, r2, r3)
        b13 = max(a13, a12 + a23)
        a123, a122 = a1 + b13 + a3, a1 + a12 + a2
        a = max(a123, a122)
        reused = pos(pos(a - a122) * ((1 - 2 * r3) ** 2 - la * la) / 2)
        d.check(area(r1) + area(r2) + area(r3), ratio * a - reused)
    elif f == MIDDLE:  # middle_configuration.cu:86-154
        for c12, c13, c23 in ((inner_outer(la, r1, r2), inner_inner(la, r1, r3), outer_inner(la, r2, r3)),
                              (outer_inner(la, r1, r2), outer_outer(r1, r3), inner_outer(la, r2, r3))):
            if d.gate(c12 + c23 - c13) >= 0:  # r2 stops r3: the piece between r1 and r2
                d.check((area(r1) + area(r2)) / 2, ratio * c12)
            else:  # r1 stops r3: the piece between r1 and r3 with all of r2
                d.check((area(r1) + area(r3)) / 2 + area(r2), ratio * c13)
    elif f == END_2D:  # end_configuration.cu:71-101
        disk = area(r2) + area(r1) / 2
        a = inner_outer(la, r1, r2) + s_out(r2)  # r1 inside
        reuse = reuse_angle(a - s_in(la, r1)) * pos((1 - 2 * r2) ** 2 - la * la) / 2
        d.check(disk, ratio * a - reuse)
        a = outer_inner(la, r1, r2) + s_in(la, r2)  # r1 outside
        reuse = reuse_angle(a - s_out(r1)) * pos(1 - (la + 2 * r2) ** 2) / 2
        d.check(disk, ratio * a - reuse)
    elif f == END_3D:  # end_configuration.cu:154-210
        disk = area(r2) + area(r3) + area(r1) / 2
        # r1 inside
        a12, a23, a13 = inner_outer(la, r1, r2), outer_inner(la, r2, r3), inner_inner(la, r1, r3)
        if d.gate(a12 + a23 - a13) < 0:  # otherwise handled as two 2-disk pieces
            a = a13 + s_in(la, r3)
            reuse = reuse_angle(a - (a12 + s_out(r2))) * pos(1 - (la + 2 * r3) ** 2) / 2
            d.check(disk, ratio * a - reuse)
        # r1 outside
        a12, a23, a13 = outer_inner(la, r1, r2), inner_outer(la, r2, r3), outer_outer(r1, r3)
        if d.gate(a12 + a23 - a13) < 0:
            a = a13 + s_out(r3)
            reuse = reuse_angle(a - (a12 + s_in(la, r2))) * pos((1 - 2 * r3) ** 2 - la * la) / 2
            d.check(disk, ratio * a - reuse)
    else:
        raise ValueError(f)
    return d.holds, min(d.margins), d.pairs
