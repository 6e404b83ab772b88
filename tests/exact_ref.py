"""Exact-rational restatement of the reference's lane-packing criteria — TEST INFRASTRUCTURE.

An independent second restatement, used to pin the C++ oracle (oracle/cpl_oracle.hpp).
It is written from the reference's own sources, not from the oracle:

* interval arithmetic: code/algcuda/cuda-utils/cuh/include/algcuda/interval.cuh
  (operators :30-355, asin widening :420-438, interval asin :617-624) and
  code/algcuda/cuda-utils/cpp/include/algcuda/interval.hpp (Uncertain<bool>
  :32-82, comparisons :134-170, :275-289, tighten/join/abs :186-232);
* geometry: code/lane_packing/collision_angles.cu;
* criteria and kernel bodies: code/lane_packing/start_configuration.cu,
  middle_configuration.cu, end_configuration.cu;
* grid rule: code/subintervals.hpp:31-42; ordered minimum:
  code/algcuda/cuda-utils/cpp/include/algcuda/atomic_float.hpp:56-66.

Every directed operation of the reference (__dadd_rd, __dmul_ru, __ddiv_rd, ...)
is evaluated exactly in fractions.Fraction and then rounded toward -inf or +inf
to a double, with IEEE 754's signed-zero rules.  Round-to-nearest operations of
the reference (get_subinterval's arithmetic, the single-placement bisection's
midpoint) are Python float operations, which are IEEE round-to-nearest; the one
multiply-add nvcc contracts in get_subinterval is an exact product-sum rounded
once.

The only code shared with the oracle is the round-to-nearest arcsine: the
reference calls CUDA's closed-source ::asin (interval.cuh:421,431), for which
the project's own asin_rn stands in on every side; `asin_rn` is injected by the
caller (the tests pass the oracle library's, whose accuracy is pinned against
mpmath separately).  Everything around it (the five-ulp widening, clamps,
domain trap) is restated here.

Verdict bytes use the encoding of include/cpl.h: bits 0-2 verdict, bit 3 the
start family's "possibly does not fit" flag, bits 4-7 the trap code (TRAP) or
the 1-based failing criterion (CRITICAL).  Traps are raised as exceptions and
reported as the first trap in the reference's evaluation order.
"""
import math
import sys
from fractions import Fraction

DBL_MAX = sys.float_info.max
INF = math.inf

# code/lane_packing/values.hpp:28-32; interval.cuh:358-365
CRITICAL_RATIO = 0.5642
LB_PI = 3.141592653589793
LAMBDA_MIN = 0.5
LAMBDA_MAX = 0.9900000000000001
UB_PI_TWO = 1.5707963267948968

START, START_END_2D, START_END_3D, MIDDLE, END_2D, END_3D = range(6)
THREE_D = {START, START_END_3D, MIDDLE, END_3D}
PROVEN, CRITICAL, NOFIT, EMPTY, OOD, TRAP = range(6)
# trap codes of include/cpl.h
T_ASIN, T_OBT_X, T_OBT_Q, T_ACU_X = 1, 2, 3, 4


class Trap(Exception):
    """A device consistency check of the reference that calls trap()."""

    def __init__(self, code):
        super().__init__(code)
        self.code = code


# ---------------------------------------------------------------------------
# Directed rounding, exact
# ---------------------------------------------------------------------------
def _round(q, up):
    """The rational q rounded toward +inf (up) or -inf to a double (q != 0)."""
    try:
        f = q.numerator / q.denominator  # correctly rounded to nearest
    except OverflowError:
        f = INF if q > 0 else -INF
    if math.isinf(f):
        if f > 0:
            return INF if up else DBL_MAX
        return -DBL_MAX if up else -INF
    e = Fraction(f)
    if up and e < q:
        return math.nextafter(f, INF)
    if not up and e > q:
        return math.nextafter(f, -INF)
    return f


def _finite(*xs):
    return all(math.isfinite(x) for x in xs)


def dadd(a, b, up):
    """__dadd_rd (up=False) / __dadd_ru (up=True)."""
    if not _finite(a, b):
        return a + b
    q = Fraction(a) + Fraction(b)
    if q == 0:
        if a == 0 and b == 0 and math.copysign(1.0, a) == math.copysign(1.0, b):
            return a
        return 0.0 if up else -0.0
    return _round(q, up)


def dmul(a, b, up):
    """__dmul_rd / __dmul_ru."""
    if not _finite(a, b):
        return a * b
    if a == 0 or b == 0:
        return math.copysign(0.0, math.copysign(1.0, a) * math.copysign(1.0, b))
    return _round(Fraction(a) * Fraction(b), up)


def ddiv(a, b, up):
    """__ddiv_rd / __ddiv_ru (b != 0 on every call of the path)."""
    if not _finite(a, b):
        return a / b
    if a == 0:
        return math.copysign(0.0, math.copysign(1.0, a) * math.copysign(1.0, b))
    return _round(Fraction(a) / Fraction(b), up)


def fma_rn(a, b, c):
    """a * b + c rounded once to nearest (nvcc's contraction of lb + diff * i)."""
    q = Fraction(a) * Fraction(b) + Fraction(c)
    if q == 0:
        return 0.0
    return q.numerator / q.denominator


RD, RU = False, True


# ---------------------------------------------------------------------------
# Uncertain<bool> (interval.hpp:32-82): (definitely, possibly)
# ---------------------------------------------------------------------------
def definitely(u):
    return u[0]


def possibly(u):
    return u[1]


def u_not(u):
    return (not u[1], not u[0])


# ---------------------------------------------------------------------------
# Interval<double>
# ---------------------------------------------------------------------------
class IV:
    __slots__ = ("lb", "ub")

    def __init__(self, lb, ub=None):
        self.lb = lb
        self.ub = lb if ub is None else ub

    def copy(self):
        return IV(self.lb, self.ub)

    def __repr__(self):
        return f"IV({self.lb!r}, {self.ub!r})"

    # interval.cuh:38-104: a + b, a - b; scalars become point intervals
    def __add__(self, o):
        o = o if isinstance(o, IV) else IV(o)
        return IV(dadd(self.lb, o.lb, RD), dadd(self.ub, o.ub, RU))

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, IV) else IV(o)
        return IV(dadd(self.lb, -o.ub, RD), dadd(self.ub, -o.lb, RU))

    def __rsub__(self, s):
        return IV(s) - self

    # interval.cuh:185-197: scalar * interval.  For n >= 0 the reference
    # computes BOTH bounds with mul_rd; restated as written.
    def scaled(self, n):
        if n >= 0:
            return IV(dmul(self.lb, n, RD), dmul(self.ub, n, RD))
        return IV(dmul(self.ub, n, RD), dmul(self.lb, n, RU))

    def __rmul__(self, n):
        return self.scaled(n)

    # interval.cuh:277-316
    def __truediv__(self, o):
        if o.lb > 0:
            if self.lb >= 0:
                return IV(ddiv(self.lb, o.ub, RD), ddiv(self.ub, o.lb, RU))
            if self.ub <= 0:
                return IV(ddiv(self.lb, o.lb, RD), ddiv(self.ub, o.ub, RU))
            return IV(ddiv(self.lb, o.lb, RD), ddiv(self.ub, o.lb, RU))
        if o.ub < 0:
            if self.lb >= 0:
                return IV(ddiv(self.ub, o.ub, RD), ddiv(self.lb, o.lb, RU))
            if self.ub <= 0:
                return IV(ddiv(self.ub, o.lb, RD), ddiv(self.lb, o.ub, RU))
            return IV(ddiv(self.ub, o.ub, RD), ddiv(self.lb, o.ub, RU))
        return IV(-INF, INF)  # DBL_MAX * DBL_MAX, interval.cuh:241-243

    # interval.hpp:216-232
    def abs(self):
        l = u = 0.0
        if self.ub > 0:
            u = self.ub
        else:
            l = -self.ub
        if self.lb < 0:
            u = -self.lb if u < -self.lb else u
        else:
            l = self.lb
        return IV(l, u)

    # interval.cuh:352-355
    def square(self):
        a = self.abs()
        return IV(dmul(a.lb, a.lb, RD), dmul(a.ub, a.ub, RU))

    # interval.hpp:186-214
    def tighten_lb(self, v):
        if self.lb < v:
            self.lb = v

    def tighten_ub(self, v):
        if v < self.ub:
            self.ub = v

    def do_join(self, o):
        if o.lb < self.lb:
            self.lb = o.lb
        if o.ub > self.ub:
            self.ub = o.ub
        return self

    def empty(self):
        return self.lb > self.ub

    # interval.hpp:134-170, 275-289
    def __lt__(self, o):
        return (self.ub < o.lb, self.lb < o.ub)

    def __le__(self, o):
        return (self.ub <= o.lb, self.lb <= o.ub)

    def __gt__(self, o):
        return o.__lt__(self)

    def __ge__(self, o):
        return o.__le__(self)


class Reference:
    """The reference's arithmetic and criteria over an injected round-to-nearest arcsine.

    asin_rn: a function double -> double (the project's asin_rn)."""

    def __init__(self, asin_rn):
        self._asin_rn = asin_rn
        self._asin_memo = {}
        self._lbrsp_memo = {}
        self.record = None  # a list: receives (lower bound of disk area, bound) of every area comparison

    def _holds(self, lb_disk_area, ub_bound):
        """The area comparison lb_disk_area >= ub_bound of every criterion."""
        if self.record is not None:
            self.record.append((lb_disk_area, ub_bound))
        return lb_disk_area >= ub_bound

    def asin_rn(self, x):
        y = self._asin_memo.get(x)
        if y is None:
            y = self._asin_memo[x] = self._asin_rn(x)
        return y

    # interval.cuh:420-438
    def dasin_rd(self, x):
        y = self.asin_rn(x)
        for _ in range(5):
            y = math.nextafter(y, -2.0)
        if y < -UB_PI_TWO:
            y = -UB_PI_TWO
        if x >= 0.0 and y < 0.0:
            y = 0.0
        return y

    def dasin_ru(self, x):
        y = self.asin_rn(x)
        for _ in range(5):
            y = math.nextafter(y, 2.0)
        if y > UB_PI_TWO:
            y = UB_PI_TWO
        if x <= 0.0 and y > 0.0:
            y = 0.0
        return y

    # interval.cuh:617-624
    def asin(self, x):
        if x.lb < -1.0 or x.ub > 1.0:
            raise Trap(T_ASIN)
        return IV(self.dasin_rd(x.lb), self.dasin_ru(x.ub))

    # ---- collision_angles.cu -------------------------------------------------
    @staticmethod
    def is_obtuse(hypot, s1, s2):  # :29-31
        return hypot.square() > s1.square() + s2.square()

    def compute_collision_angle(self, lc, sc, ab):  # :37-117
        lc, sc = lc.copy(), sc.copy()
        lc.tighten_lb(sc.lb)
        sc.tighten_ub(lc.ub)
        result = IV(0.0, -1.0)
        obt = self.is_obtuse(lc, sc, ab)
        lc_sq = lc.square()
        sc_sq = sc.square()
        add_term = (lc_sq - sc_sq) / ab
        add_term.tighten_lb(0.0)
        if possibly(obt):
            lx = 0.5 * (add_term + ab)
            sx = 0.5 * (add_term - ab)
            sx.tighten_lb(0.0)
            sx.tighten_ub(sc.ub)
            lx.tighten_ub(lc.ub)
            if sx.empty() or lx.empty():
                if definitely(obt):
                    raise Trap(T_OBT_X)
            else:
                lquot = lx / lc
                lquot.tighten_ub(1.0)
                squot = sx / sc
                squot.tighten_ub(1.0)
                if lquot.empty() or squot.empty():
                    if definitely(obt):
                        raise Trap(T_OBT_Q)
                else:
                    alpha = self.asin(lquot)
                    beta = self.asin(squot)
                    result = alpha - beta
                    result.tighten_lb(0.0)
        if not definitely(obt):
            lx = 0.5 * (ab + add_term)
            sx = 0.5 * (ab - add_term)
            sx.tighten_lb(0.0)
            if sx.empty():
                if not possibly(obt):
                    raise Trap(T_ACU_X)
            else:
                alpha = self.asin(lx / lc)
                beta = self.asin(sx / sc)
                beta.tighten_lb(0.0)
                r = alpha + beta
                if result.empty():
                    result = r
                else:
                    result.do_join(r)
        return result

    def collision_angle_inner_inner(self, la, ra, rb):  # :119-142
        lc = la.lb + IV(ra.ub, ra.ub)
        sc = la.lb + IV(rb.ub, rb.ub)
        ab = IV(ra.ub, ra.ub) + rb.ub
        ub = self.compute_collision_angle(lc, sc, ab).ub
        lc = la.ub + IV(ra.lb, ra.lb)
        sc = la.ub + IV(rb.lb, rb.lb)
        ab = IV(ra.lb, ra.lb) + rb.lb
        lb = 0.0 if ab.lb <= 0.0 else self.compute_collision_angle(lc, sc, ab).lb
        return IV(lb, ub)

    def _pick_longer(self, ca, cb, ab):  # the shared body of :144-159 and :176-191
        ca_longer = cb <= ca
        if definitely(ca_longer):
            return self.compute_collision_angle(ca, cb, ab)
        if definitely(u_not(ca_longer)):
            return self.compute_collision_angle(cb, ca, ab)
        result = self.compute_collision_angle(ca, cb, ab)
        result.do_join(self.compute_collision_angle(cb, ca, ab))
        return result

    def do_collision_angle_outer_inner(self, la, ra, rb):  # :144-159
        return self._pick_longer(1.0 - ra, la + rb, ra + rb)

    def do_collision_angle_inner_outer(self, la, ra, rb):  # :176-191
        return self._pick_longer(la + ra, 1.0 - rb, ra + rb)

    def collision_angle_outer_inner(self, la, ra, rb):  # :161-174
        if possibly((2.0 * (ra + rb)) <= (1.0 - la)):
            lb = 0.0
        else:
            lb = self.do_collision_angle_outer_inner(la, IV(ra.lb), IV(rb.lb)).lb
        ub = self.do_collision_angle_outer_inner(la, IV(ra.ub), IV(rb.ub)).ub
        return IV(lb, ub)

    def collision_angle_inner_outer(self, la, ra, rb):  # :193-205
        if possibly((2.0 * (ra + rb)) <= (1.0 - la)):
            lb = 0.0
        else:
            lb = self.do_collision_angle_inner_outer(la, IV(ra.lb), IV(rb.lb)).lb
        ub = self.do_collision_angle_inner_outer(la, IV(ra.ub), IV(rb.ub)).ub
        return IV(lb, ub)

    def collision_angle_outer_outer(self, ra, rb):  # :207-231
        lc = IV(1.0, 1.0) - rb.ub
        sc = IV(1.0, 1.0) - ra.ub
        ab = IV(ra.ub, ra.ub) + rb.ub
        ub = self.compute_collision_angle(lc, sc, ab).ub
        lc = IV(1.0, 1.0) - rb.lb
        sc = IV(1.0, 1.0) - ra.lb
        ab = IV(ra.lb, ra.lb) + rb.lb
        lb = 0.0 if ab.lb <= 0.0 else self.compute_collision_angle(lc, sc, ab).lb
        return IV(lb, ub)

    def semicircle_angle_outer(self, r):  # :233-238
        return IV(self.dasin_rd(ddiv(r.lb, dadd(1.0, -r.lb, RU), RD)),
                  self.dasin_ru(ddiv(r.ub, dadd(1.0, -r.ub, RD), RU)))

    def semicircle_angle_inner(self, la, r):  # :240-246
        return IV(self.dasin_rd(ddiv(r.lb, dadd(la.ub, r.lb, RU), RD)),
                  self.dasin_ru(ddiv(r.ub, dadd(la.lb, r.ub, RD), RU)))

    # ---- lower_bound_radius_single_placement.cu:26-61 -------------------------
    def lower_bound_radius_single_placement(self, lb_lambda):
        if lb_lambda in self._lbrsp_memo:
            return self._lbrsp_memo[lb_lambda]
        ub_rhs_factor = ddiv(CRITICAL_RATIO, LB_PI, RU)
        ub_rhs = dmul(ub_rhs_factor, dadd(1.0, -dmul(lb_lambda, lb_lambda, RD), RU), RU)
        ub, lb = 0.25, 5.0e-324
        while True:
            mid = 0.5 * (lb + ub)
            if mid <= lb or mid >= ub:
                break
            r_sq = dmul(mid, mid, RD)
            ub_quot = ddiv(mid, dadd(1.0, -mid, RD), RU)
            lb_lhs = ddiv(r_sq, self.dasin_ru(ub_quot), RD)
            if lb_lhs >= ub_rhs:
                ub = mid
            else:
                lb = mid
        self._lbrsp_memo[lb_lambda] = ub
        return ub

    # ---- admissible ranges ------------------------------------------------------
    def compute_r1_range(self, la):  # start :58-66 = middle :66-74 = end :55-63
        by_single = self.lower_bound_radius_single_placement(la.lb)
        ub_width = (0.5 * (1.0 - la)).ub
        lb_width = (0.25 * (1.0 - la)).lb
        return IV(lb_width, ub_width if ub_width < by_single else by_single)

    @staticmethod
    def _end_lane_lb(la, r_ub):  # end :65-69, :148-152
        v = dmul(0.5, dadd(1.0, -dadd(la.ub, dmul(2.0, r_ub, RU), RU), RD), RD)
        return 0.0 if v < 0.0 else v

    def compute_r2_range(self, f, la, r1):
        if f == START_END_2D:  # start :74-77
            return IV((0.5 * (1.0 - la - 2.0 * r1)).lb, r1.ub)
        if f in (END_2D, END_3D):  # end :65-69
            return IV(self._end_lane_lb(la, r1.ub), r1.ub)
        return IV((0.25 * (1.0 - la)).lb, r1.ub)  # start :68-71, middle :76-79

    def compute_r3_range(self, f, la, r1, r2):
        if f == START_END_3D:  # start :84-87
            return IV((0.5 * (1.0 - la - 2.0 * r2)).lb, r2.ub)
        if f == END_3D:  # end :148-152
            return IV(self._end_lane_lb(la, r2.ub), r2.ub)
        return IV((0.25 * (1.0 - la)).lb, r2.ub)  # start :79-82, middle :81-84

    # ---- criteria -------------------------------------------------------------------
    @staticmethod
    def _ub_ratio(la):
        return dadd(1.0, -dmul(la.lb, la.lb, RD), RU)

    @staticmethod
    def _lb_area(r):
        return dmul(LB_PI, dmul(r.lb, r.lb, RD), RD)

    def start_case_1(self, la, r1, r2, a12):  # start :93-101
        ratio = self._ub_ratio(la)
        a1, a2 = self._lb_area(r1), self._lb_area(r2)
        a = self.semicircle_angle_outer(r1) + a12
        used = dmul(ratio, dmul(0.5, a.ub, RU), RU)
        disk = dadd(a1, dmul(0.5, a2, RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, used, RU))

    def start_case_2(self, la, r1, r2, r3, a13):  # start :103-112
        ratio = self._ub_ratio(la)
        a1, a2, a3 = self._lb_area(r1), self._lb_area(r2), self._lb_area(r3)
        a = self.semicircle_angle_outer(r1) + a13
        used = dmul(ratio, dmul(0.5, a.ub, RU), RU)
        disk = dadd(dadd(a1, a2, RD), dmul(0.5, a3, RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, used, RU))

    def start_end_2d(self, la, r1, r2):  # start :169-188
        a1 = self.semicircle_angle_outer(r1)
        a2 = self.semicircle_angle_inner(la, r2)
        a12 = self.collision_angle_outer_inner(la, r1, r2)
        total = a1 + a2 + a12
        reuse = a12 + a2 - a1
        reuse.tighten_lb(0.0)
        used = dmul(self._ub_ratio(la), dmul(0.5, total.ub, RU), RU)
        rho = 1.0 - (la + 2.0 * r2).square()
        rho.tighten_lb(0.0)
        reusable = dmul(dmul(0.5, reuse.lb, RD), rho.lb, RD)
        disk = dadd(self._lb_area(r2), self._lb_area(r1), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, dadd(used, -reusable, RU), RU))

    def start_end_3d(self, la, r1, r2, r3):  # start :227-265
        a1 = self.semicircle_angle_outer(r1)
        a2 = self.semicircle_angle_inner(la, r2)
        a3 = self.semicircle_angle_outer(r3)
        a12 = self.collision_angle_outer_inner(la, r1, r2)
        a13 = self.collision_angle_outer_outer(r1, r2)  # :232 passes r2, restated as written
        a23 = self.collision_angle_inner_outer(la, r2, r3)
        b13_123 = a12 + a23
        b13 = IV(b13_123.lb if a13.lb < b13_123.lb else a13.lb, b13_123.ub if a13.ub < b13_123.ub else a13.ub)
        a123 = a1 + b13 + a3
        a122 = a1 + a12 + a2
        a = IV(a122.lb if a123.lb < a122.lb else a123.lb, a122.ub if a123.ub < a122.ub else a123.ub)
        reusable_angle = (a - a122).lb
        if reusable_angle < 0.0:
            reusable_angle = 0.0
        outer_rad = dadd(1.0, -dmul(2.0, r3.ub, RU), RD)
        outer_sq = dmul(outer_rad, outer_rad, RD)
        inner_sq = dmul(la.ub, la.ub, RU)
        reused = dmul(dmul(0.5, reusable_angle, RD), dadd(outer_sq, -inner_sq, RD), RD)
        if reused < 0.0:
            reused = 0.0
        used = dmul(self._ub_ratio(la), dmul(0.5, a.ub, RU), RU)
        used = dadd(used, -reused, RU)
        disk = dadd(self._lb_area(r1), dadd(self._lb_area(r2), self._lb_area(r3), RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, used, RU))

    def _middle_tail(self, la, r1, r2, r3, c12, c13, c23):  # middle :93-118 = :127-152
        ratio = self._ub_ratio(la)
        case = (c12 + c23) >= c13
        a1, a2 = self._lb_area(r1), self._lb_area(r2)
        if possibly(case):
            used = dmul(ratio, dmul(0.5, c12.ub, RU), RU)
            disk = dmul(0.5, dadd(a1, a2, RD), RD)
            if not self._holds(disk, dmul(CRITICAL_RATIO, used, RU)):
                return False
        if possibly(u_not(case)):
            used = dmul(ratio, dmul(0.5, c13.ub, RU), RU)
            a3 = self._lb_area(r3)
            disk = dadd(dmul(0.5, dadd(a1, a3, RD), RD), a2, RD)
            if not self._holds(disk, dmul(CRITICAL_RATIO, used, RU)):
                return False
        return True

    def middle_case_1_1(self, la, r1, r2, r3):  # middle :86-120
        c12 = self.collision_angle_inner_outer(la, r1, r2)
        c13 = self.collision_angle_inner_inner(la, r1, r3)
        c23 = self.collision_angle_outer_inner(la, r2, r3)
        return self._middle_tail(la, r1, r2, r3, c12, c13, c23)

    def middle_case_2_1(self, la, r1, r2, r3):  # middle :122-154
        c12 = self.collision_angle_outer_inner(la, r1, r2)
        c13 = self.collision_angle_outer_outer(r1, r3)
        c23 = self.collision_angle_inner_outer(la, r2, r3)
        return self._middle_tail(la, r1, r2, r3, c12, c13, c23)

    def end_inner_outer_2d(self, la, r1, r2):  # end :71-85
        a = self.collision_angle_inner_outer(la, r1, r2) + self.semicircle_angle_outer(r2)
        used = dmul(self._ub_ratio(la), dmul(0.5, a.ub, RU), RU)
        rho = (1.0 - 2.0 * r2).square() - la.square()
        rho.tighten_lb(0.0)
        alpha = (a - self.semicircle_angle_inner(la, r1)).lb
        reusable = dmul(dmul(0.5, alpha, RD), rho.lb, RD)
        disk = dadd(self._lb_area(r2), dmul(0.5, self._lb_area(r1), RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, dadd(used, -reusable, RU), RU))

    def end_outer_inner_2d(self, la, r1, r2):  # end :87-101
        a = self.collision_angle_outer_inner(la, r1, r2) + self.semicircle_angle_inner(la, r2)
        used = dmul(self._ub_ratio(la), dmul(0.5, a.ub, RU), RU)
        rho = 1.0 - (la + 2.0 * r2).square()
        rho.tighten_lb(0.0)
        alpha = (a - self.semicircle_angle_outer(r1)).lb
        reusable = dmul(dmul(0.5, alpha, RD), rho.lb, RD)
        disk = dadd(self._lb_area(r2), dmul(0.5, self._lb_area(r1), RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, dadd(used, -reusable, RU), RU))

    def _end_3d_tail(self, la, r1, r2, r3, a, a12, a2, rho):  # end :166-180 = :195-209
        alpha = a - (a12 + a2)
        rho.tighten_lb(0.0)
        reusable = dmul(dmul(0.5, alpha.lb, RD), rho.lb, RD)
        used = dmul(self._ub_ratio(la), dmul(0.5, a.ub, RU), RU)
        disk = dadd(dadd(self._lb_area(r2), self._lb_area(r3), RD), dmul(0.5, self._lb_area(r1), RD), RD)
        return self._holds(disk, dmul(CRITICAL_RATIO, dadd(used, -reusable, RU), RU))

    def end_r1_inner_3d(self, la, r1, r2, r3):  # end :154-181
        a12 = self.collision_angle_inner_outer(la, r1, r2)
        a23 = self.collision_angle_outer_inner(la, r2, r3)
        a13 = self.collision_angle_inner_inner(la, r1, r3)
        if definitely((a12 + a23) >= a13):
            return True
        a3 = self.semicircle_angle_inner(la, r3)
        a = a13 + a3
        a2 = self.semicircle_angle_outer(r2)
        rho = 1.0 - (la + 2.0 * r3).square()
        return self._end_3d_tail(la, r1, r2, r3, a, a12, a2, rho)

    def end_r1_outer_3d(self, la, r1, r2, r3):  # end :183-210
        a12 = self.collision_angle_outer_inner(la, r1, r2)
        a23 = self.collision_angle_inner_outer(la, r2, r3)
        a13 = self.collision_angle_outer_outer(r1, r3)
        if definitely((a12 + a23) >= a13):
            return True
        a3 = self.semicircle_angle_outer(r3)
        a = a13 + a3
        a2 = self.semicircle_angle_inner(la, r2)
        rho = (1.0 - 2.0 * r3).square() - la.square()
        return self._end_3d_tail(la, r1, r2, r3, a, a12, a2, rho)

    # ---- kernel bodies ------------------------------------------------------------
    @staticmethod
    def _no_fit(la, r1):  # start :131, middle :173, end :115, :236
        return dmul(2.0, r1.ub, RU) >= dadd(1.0, -la.ub, RD)

    def failing_criteria(self, f, la, r1, r2, r3, all_criteria=False):
        """1-based indices of the criteria of family f that fail on the tightened
        cell, in the reference's order; stops at the first unless all_criteria
        (the output_criticals mode, which keeps evaluating).  Raises Trap."""
        fails = []

        def note(k):
            fails.append(k)
            return not all_criteria

        if f == START:  # start :137-162
            a12 = self.collision_angle_outer_inner(la, r1, r2)
            a13 = self.collision_angle_outer_outer(r1, r3)
            a23 = self.collision_angle_inner_outer(la, r2, r3)
            case1 = (a12 + a23) > a13
            if possibly(case1) and not self.start_case_1(la, r1, r2, a12):
                if note(1):
                    return fails
            if possibly(u_not(case1)) and not self.start_case_2(la, r1, r2, r3, a13):
                note(2)
        elif f == START_END_2D:  # start :202
            if not self.start_end_2d(la, r1, r2):
                note(1)
        elif f == START_END_3D:  # start :284
            if not self.start_end_3d(la, r1, r2, r3):
                note(1)
        else:
            first, second = {
                MIDDLE: (self.middle_case_1_1, self.middle_case_2_1),      # middle :180, :191
                END_2D: (self.end_inner_outer_2d, self.end_outer_inner_2d),  # end :122, :133
                END_3D: (self.end_r1_inner_3d, self.end_r1_outer_3d),        # end :243, :254
            }[f]
            args = (la, r1, r2) if f == END_2D else (la, r1, r2, r3)
            if not first(*args):
                if note(1):
                    return fails
            if not second(*args):
                note(2)
        return fails

    def kernel_body(self, f, la, r1_, r2_, r3):
        """Verdict byte of one reference-grid cell (la, r1_, r2_, r3 before
        tightening), following the kernel bodies (start :126-162, :199-211,
        :279-293; middle :166-201; end :107-143, :225-264).  EMPTY for every
        family follows the library's verdict definition (the reference checks
        emptiness in end 3d only, end :232-234; elsewhere a tightened interval of
        the reference grids cannot be empty)."""
        r2 = r2_.copy()
        if f in THREE_D:
            r2.tighten_lb(r3.lb)
        r1 = r1_.copy()
        r1.tighten_lb(r2.lb)
        if r1.empty() or r2.empty():
            return EMPTY
        nofit_flag = 0
        if f in (START, MIDDLE, END_2D, END_3D) and self._no_fit(la, r1):
            if f != START:
                return NOFIT  # middle :173-178, end :115-120, :236-241 trap here
            nofit_flag = 8  # start :131-135 prints and continues
        try:
            fails = self.failing_criteria(f, la, r1, r2, r3)
        except Trap as t:
            return TRAP | nofit_flag | (t.code << 4)
        if fails:
            return CRITICAL | nofit_flag | (fails[0] << 4)
        return PROVEN | nofit_flag

    def kernel_body_all(self, f, la, r1_, r2_, r3):
        """The output_criticals form of the kernel body (every criterion is
        evaluated, one printed line per failure): -> (early verdict or None,
        start no-fit flag, failing criteria, trap code or 0, tightened r1, r2).
        The start family's no-fit line is printed before its criteria run
        (start :131-135), so it stands when the cell then traps."""
        r2 = r2_.copy()
        if f in THREE_D:
            r2.tighten_lb(r3.lb)
        r1 = r1_.copy()
        r1.tighten_lb(r2.lb)
        if r1.empty() or r2.empty():
            return EMPTY, False, [], 0, r1, r2
        nofit = False
        if f in (START, MIDDLE, END_2D, END_3D) and self._no_fit(la, r1):
            if f != START:
                return NOFIT, False, [], 0, r1, r2
            nofit = True
        try:
            return None, nofit, self.failing_criteria(f, la, r1, r2, r3, all_criteria=True), 0, r1, r2
        except Trap as t:
            return None, nofit, [], t.code, r1, r2

    def area_terms(self, f, la, r1_, r2_, r3):
        """The area comparisons of the output_criticals evaluation of a cell, in
        evaluation order, or None when the cell is decided before the criteria
        or traps."""
        self.record = []
        try:
            early, _, _, code, _, _ = self.kernel_body_all(f, la, r1_, r2_, r3)
            return None if early is not None or code else self.record
        finally:
            self.record = None

    # ---- synthetic cells (the 4-parameter refinement workload) -----------------------
    @staticmethod
    def r3_from_position(rng, p):
        """r3 = the position interval p in [0, 1] mapped onto the admissible r3
        range; for p = [k/n, (k+1)/n] this is get_subinterval(rng, k, n)."""
        w = rng.ub - rng.lb
        lo = fma_rn(w, p.lb, rng.lb)
        hi = rng.ub if p.ub == 1.0 else fma_rn(w, p.ub, rng.lb)
        return IV(lo, hi)

    def synthetic(self, f, la, r1, r2, p3, clip=True):
        """Verdict byte of a synthetic cell: radii clipped to the family's
        admissible ranges (OOD when a clip is empty), r3 from its position."""
        if clip:
            rr = self.compute_r1_range(la)
            r1 = IV(max(r1.lb, rr.lb), min(r1.ub, rr.ub))
            if r1.empty():
                return OOD
            rr = self.compute_r2_range(f, la, r1)
            r2 = IV(max(r2.lb, rr.lb), min(r2.ub, rr.ub))
            if r2.empty():
                return OOD
        r3 = IV(0.0, 0.0)
        if f in THREE_D:
            rng = self.compute_r3_range(f, la, r1, r2)
            if rng.empty():
                return OOD
            r3 = self.r3_from_position(rng, p3)
        return self.kernel_body(f, la, r1, r2, r3)


# ---- subintervals.hpp:31-42 --------------------------------------------------------
def get_subinterval(iv, i, n):
    lb, ub = iv.lb, iv.ub
    diff = ub - lb
    diff /= n
    l = fma_rn(diff, float(i), lb)
    u = fma_rn(diff, float(i + 1), lb)
    if i + 1 == n:
        u = ub
    return IV(l, u)


# ---- atomic_float.hpp:56-66 -------------------------------------------------------
def ordered_from_bits(b):
    """Atomic_ordered_double::transform on the 64-bit pattern b."""
    mask = ((-(b >> 63)) & 0xFFFFFFFFFFFFFFFF) | (1 << 63)
    return b ^ mask


def bits_from_ordered(v):
    """Atomic_ordered_double::inverse_transform."""
    mask = (((v >> 63) - 1) & 0xFFFFFFFFFFFFFFFF) | (1 << 63)
    return v ^ mask


# ---- one reference-grid launch (the six kernels' loops) -----------------------------
def mode_r(ref, f, outer, subdiv, output_criticals=False):
    """Every cell of the reference grid of family f over the lambda interval
    outer, in the kernels' nesting (la, r1, r2, r3; start :115-125, middle
    :157-166, end :104-109, :213-223; end 3d skips empty r3 ranges :222-224).

    first = min la.lb of a cell with a failing criterion.  A reference thread
    walks its cells in increasing la and returns at its first failure, so every
    cell with la.lb < first is reached: traps and the fit-check traps of the
    middle and end families count there (everywhere with output_criticals,
    which evaluates every cell).  rc: -3 trap (code: smallest trapping la.lb,
    then smallest code), -4 fit check, 0 otherwise.  records: the printed lines
    as (kind, la, r1, r2, r3) tuples, kind 1/2 = failing criterion, 3 = the
    start family's "possibly does not fit" line."""
    nla, n1, n2, n3 = subdiv
    if f not in THREE_D:
        n3 = 1
    first = DBL_MAX
    fail = {}  # kind (0 fit check, c trap code) -> smallest la.lb
    evaluated = nofit_flags = 0
    records = []
    for ia in range(nla):
        la = get_subinterval(outer, ia, nla)
        r1_range = ref.compute_r1_range(la)
        for i1 in range(n1):
            r1_ = get_subinterval(r1_range, i1, n1)
            r2_range = ref.compute_r2_range(f, la, r1_)
            for i2 in range(n2):
                r2_ = get_subinterval(r2_range, i2, n2)
                if f in THREE_D:
                    r3_range = ref.compute_r3_range(f, la, r1_, r2_)
                    if f == END_3D and r3_range.empty():
                        continue
                for i3 in range(n3):
                    r3 = get_subinterval(r3_range, i3, n3) if f in THREE_D else IV(0.0, 0.0)
                    evaluated += 1
                    if not output_criticals:
                        vb = ref.kernel_body(f, la, r1_, r2_, r3)
                        v = vb & 7
                        nofit_flags += (vb >> 3) & 1
                        if v == CRITICAL:
                            first = min(first, la.lb)
                        elif v in (NOFIT, TRAP):
                            k = 0 if v == NOFIT else vb >> 4
                            fail[k] = min(fail.get(k, INF), la.lb)
                        continue
                    early, nofit, fails, code, r1, r2 = ref.kernel_body_all(f, la, r1_, r2_, r3)
                    if code:
                        fail[code] = min(fail.get(code, INF), la.lb)
                    if early == NOFIT:
                        fail[0] = min(fail.get(0, INF), la.lb)
                    if early is not None:
                        continue
                    cell = (la.lb, la.ub, r1.lb, r1.ub, r2.lb, r2.ub, r3.lb, r3.ub)
                    if nofit:
                        nofit_flags += 1
                        records.append((3.0,) + cell)
                    if fails:
                        first = min(first, la.lb)
                    records.extend((float(k),) + cell for k in fails)
    limit = INF if output_criticals else first
    traps = sorted((v, k) for k, v in fail.items() if k and v < limit)
    rc, code = 0, 0
    if traps:
        rc, code = -3, traps[0][1]
    elif fail.get(0, INF) < limit:
        rc = -4
    return {"rc": rc, "first": first, "evaluated": evaluated, "trap_code": code, "nofit_flags": nofit_flags,
            "records": sorted(records)}
