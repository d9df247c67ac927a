"""config.txt texts of the test and benchmark workloads.

Parameter choices follow the workload table of the measurement plan:
  drop2d   2D drop, fully periodic (128x128, R=25 is the parity gate)
  rti2d    2D Rayleigh-Taylor, periodic x, walls y
  bubble3d 3D rising bubble, fully periodic
  jet      velocity inlet + free-convect outflow (2D and 3D)
"""


def _bc(sides):
    blocks = []
    for side, typ, *val in sides:
        v = f"\n    value = [{', '.join(str(x) for x in val[0])}]" if val else ""
        blocks.append(f"  {{\n    side = \"{side}\"\n    type = \"{typ}\"{v}\n  }}")
    return "boundary = [\n" + ",\n".join(blocks) + "\n]\n"


def _geom(typ, center, radius=0.0, length=None, amplitude=None):
    extra = ""
    if length is not None:
        extra += f"\n    length = [{', '.join(str(x) for x in length)}]"
    if amplitude is not None:
        extra += f"\n    amplitude = {amplitude}"
    return (f"geometry = [\n  {{\n    type = \"{typ}\"\n    center = [{', '.join(str(x) for x in center)}]"
            f"\n    radius = {radius}{extra}\n    id = 0\n  }}\n]\n")


def drop2d(n=128, radius=25.0, rho_ratio=1000, mu_ratio=100, We=5, Re=100, mobility=0.02, W=4,
           gravity=(0, 0, 0), sides=None, ny=None):
    """2D drop; sigma and mu_H derived with a |g| = 1e-6 scale, body force
    given by `gravity` (zero by default)."""
    ny = ny or n
    c = ((n + 1) / 2.0, (ny + 1) / 2.0, 0)
    sides = sides or [("x-", "periodic"), ("x+", "periodic"), ("y-", "periodic"), ("y+", "periodic")]
    return (f"case = drop\nfinalStep = 1000\nreportInterval = 100\n"
            f"N = [{n}, {ny}, 1]\nreferenceLength = {2 * radius}\n"
            f"We = {We}\nRe = {Re}\ngravity = [{', '.join(str(x) for x in gravity)}]\ngravityScale = 1e-6\n"
            f"mobility = {mobility}\ninterfaceWidth = {W}\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n\n"
            + _geom("circle", c, radius) + "\n" + _bc(sides))


def rti2d(nx=256, ny=1024, Re=256, Ca=0.26, Pe=1000, W=4, amp=0.1):
    """2D RTI: heavy fluid above the perturbed interface y0 = ny/2 + amp*nx*cos(2 pi x/nx),
    gravity -y chosen so that U = sqrt(g L) = 0.04 with L = nx."""
    g = (0.04 ** 2) / nx
    return (f"case = RTI\nfinalStep = 20000\nreportInterval = 1000\nN = [{nx}, {ny}, 1]\n"
            f"referenceLength = {nx}\nRe = {Re}\nPe = {Pe}\nCa = {Ca}\n"
            f"gravity = [0, {-g}, 0]\ninterfaceWidth = {W}\nmuRatio = 1\nrhoRatio = 3\n\n"
            + _geom("perturbation", (0, ny / 2.0, 0), 0.0, length=(nx, 0, 0), amplitude=amp) + "\n"
            + _bc([("x-", "periodic"), ("x+", "periodic"), ("y-", "wall"), ("y+", "wall")]))


def bubble3d(n=256, radius=32.0, rho_ratio=1000, mu_ratio=100, Eo=10.0, Mo=1e-3, W=4, mobility=0.02,
             g=6.25e-6, sides=None):
    """3D rising bubble, N = n + 2 per axis (ghosts inside N).  We = Eo / drho,
    Re = (Eo^3 / Mo)^(1/4); gravity -y of magnitude g."""
    N = n + 2
    drho = 1.0 - 1.0 / rho_ratio
    We = Eo / drho
    Re = (Eo ** 3 / Mo) ** 0.25
    c = (N - 1) / 2.0
    sides = sides or [(s, "periodic") for s in ("x-", "x+", "y-", "y+", "z-", "z+")]
    return (f"case = bubble\nfinalStep = 5000\nreportInterval = 500\nN = [{N}, {N}, {N}]\n"
            f"referenceLength = {2 * radius}\nWe = {We}\nRe = {Re}\ngravity = [0, {-g}, 0]\n"
            f"mobility = {mobility}\ninterfaceWidth = {W}\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n\n"
            + _geom("sphere", (c, c, c), radius) + "\n" + _bc(sides))


def jet2d(nx=48, ny=64, radius=6.0, U=0.02, rho_ratio=4, mu_ratio=2, W=4):
    """2D jet: velocity inlet on y+ through a circle of `radius` (geometry 0),
    free-convect outflow on y-, periodic x."""
    c = ((nx + 1) / 2.0, ny + 1, 0)
    return (f"case = jet\nfinalStep = 1000\nN = [{nx}, {ny}, 1]\nreferenceLength = {2 * radius}\n"
            f"referenceVelocity = {U}\nWe = 50\nRe = 200\ngravity = [0, 0, 0]\nmobility = 0.02\n"
            f"interfaceWidth = {W}\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n\n"
            + _geom("circle", c, radius) + "\n"
            + _bc([("x-", "periodic"), ("x+", "periodic"), ("y-", "free convect"), ("y+", "velocity", (0, -U, 0))]))


def jet3d(n=(24, 32, 24), radius=5.0, U=0.02, rho_ratio=4, mu_ratio=2, W=4, inlet_side="y+",
          outlet_type="free convect"):
    """3D jet: velocity inlet through a sphere-gated disc on `inlet_side`,
    outflow on the opposite side, periodic elsewhere (N includes ghosts)."""
    nx, ny, nz = n
    ax = "xyz".index(inlet_side[0])
    cen = [(nx - 1) / 2.0, (ny - 1) / 2.0, (nz - 1) / 2.0]
    cen[ax] = (n[ax] - 1) if inlet_side[1] == "+" else 0
    vel = [0, 0, 0]
    vel[ax] = -U if inlet_side[1] == "+" else U
    sides = []
    for a in "xyz":
        for sgn in "-+":
            s = a + sgn
            if s == inlet_side:
                sides.append((s, "velocity", vel))
            elif a == inlet_side[0]:
                sides.append((s, outlet_type))
            else:
                sides.append((s, "periodic"))
    return (f"case = jet\nfinalStep = 1000\nN = [{nx}, {ny}, {nz}]\nreferenceLength = {2 * radius}\n"
            f"referenceVelocity = {U}\nWe = 50\nRe = 200\ngravity = [0, 0, 0]\nmobility = 0.02\n"
            f"interfaceWidth = {W}\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n\n"
            + _geom("sphere", cen, radius) + "\n" + _bc(sides))


def box2d(sides, n=24, ny=32, case="drop", radius=6.0):
    """2D drop resting mid-box with arbitrary side conditions and gravity -y."""
    return drop2d(n=n, ny=ny, radius=radius, rho_ratio=10, mu_ratio=10, gravity=(0, -1e-5, 0), sides=sides)


def jet3d_c5(nx=192, ny=192, nz=768, radius=16.0, U=0.05, L=32.0, rho_ratio=100, mu_ratio=10, Re=500, We=50,
             modes=8, seed=0, W=4):
    """3D jet: inlet on z- through a disc of the given radius centred in x, y
    (sphere geometry on the inlet plane), u_z = U (1 + sum_m A_m cos(m theta +
    phase_m)) with A_m ~ 1e-2 N(0,1), phase_m ~ U(0, 2 pi) (seeded), convective
    outflow on z+, periodic x and y.  Interior nx x ny x nz (N includes the
    ghost ring)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    amp = 1e-2 * rng.standard_normal(modes)
    ph = rng.uniform(0.0, 2 * np.pi, modes)
    N = (nx + 2, ny + 2, nz + 2)
    cen = ((N[0] - 1) / 2.0, (N[1] - 1) / 2.0, 0)
    sides = [("x-", "periodic"), ("x+", "periodic"), ("y-", "periodic"), ("y+", "periodic"),
             ("z-", "velocity", [0, 0, U]), ("z+", "free convect")]
    return (f"case = jet\nfinalStep = 3000\nN = [{N[0]}, {N[1]}, {N[2]}]\nreferenceLength = {L}\n"
            f"referenceVelocity = {U}\nWe = {We}\nRe = {Re}\ngravity = [0, 0, 0]\nmobility = 0.02\n"
            f"interfaceWidth = {W}\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n"
            f"inletModes = {modes}\ninletModeAmplitude = [{', '.join(repr(float(a)) for a in amp)}]\n"
            f"inletModePhase = [{', '.join(repr(float(p)) for p in ph)}]\n\n"
            + _geom("sphere", cen, radius) + "\n" + _bc(sides))


def box3d(sides, n=(10, 12, 8), radius=3.0, rho_ratio=10, mu_ratio=10):
    """3D drop mid-box with arbitrary side conditions and gravity -y (N =
    interior + 2 per axis, ghosts inside N)."""
    N = tuple(x + 2 for x in n)
    c = tuple((x - 1) / 2.0 for x in N)
    return (f"case = drop\nfinalStep = 1000\nN = [{N[0]}, {N[1]}, {N[2]}]\nreferenceLength = {2 * radius}\n"
            f"We = 5\nRe = 100\ngravity = [0, -1e-5, 0]\ngravityScale = 1e-6\nmobility = 0.02\n"
            f"interfaceWidth = 4\nmuRatio = {mu_ratio}\nrhoRatio = {rho_ratio}\n\n"
            + _geom("sphere", c, radius) + "\n" + _bc(sides))


def specular(text):
    """The same configuration with slip sides reflecting specularly
    (slipWallBCMap) instead of the reference's default."""
    return "slipWall = specular\n" + text
