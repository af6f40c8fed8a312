"""tblgen on the GPU: emissivity rows of a lookup table from a monochromatic
optical-depth spectrum and a filter function.

    python -m jurassic_amd.tblgen <press> <temp> <spec> <filter>

reads the same files as the reference tool (src/tblgen.c:44-140) and prints
the same lines, "press temp u eps" with %g, for the rows it keeps:
  * <filter>: lines "nu f" (src/tblgen.c:67-74);
  * <spec>: an RFM optical-depth file; its fourth line holds
    "npts nu0 dnu nu1", the values follow, whitespace-separated
    (src/tblgen.c:77-102).
"""
import sys

import numpy as np

MAXNF = 20000        # src/tblgen.c:32
MAXNPTS = 10000000   # src/tblgen.c:35


def read_filter(path):
    """(nu, f) of the lines that hold two numbers (src/tblgen.c:67-74)."""
    nu, f = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 2:
                try:
                    a, b = float(parts[0]), float(parts[1])
                except ValueError:
                    continue
                nu.append(a)
                f.append(b)
                if len(nu) >= MAXNF:
                    raise ValueError("Too many points in filter function")
    return np.array(nu), np.array(f)


def read_spectrum(path):
    """(tau, nu0, nu1) of an RFM optical-depth file (src/tblgen.c:77-102)."""
    with open(path) as fh:
        head = [fh.readline() for _ in range(4)]
        if not head[3]:
            raise ValueError("Error while reading spectrum!")
        f4 = head[3].split()
        npts, nu0, nu1 = int(f4[0]), float(f4[1]), float(f4[3])
        if npts > MAXNPTS:
            raise ValueError("Too many points in optical depth spectrum!")
        vals = np.array(fh.read().split(), dtype=np.float64)
    # the tool integrates npts points of a zero-initialised array
    tau = np.zeros(npts)
    m = min(npts, len(vals))
    tau[:m] = vals[:m]
    return tau, nu0, nu1


def write_spectrum(path, tau, nu0, nu1):
    """An optical-depth file in the layout read_spectrum reads."""
    n = len(tau)
    with open(path, "w") as fh:
        fh.write("! optical depth spectrum\n! synthetic\n! npts nu0 dnu nu1\n")
        fh.write("%d %.10f %.10e %.10f\n" % (n, nu0, (nu1 - nu0) / (n - 1), nu1))
        for i in range(0, n, 5):
            fh.write(" ".join("%.8e" % v for v in tau[i:i + 5]) + "\n")


def rows_text(press, temp, u, eps):
    """The lines tblgen prints (src/tblgen.c:132-134)."""
    return "".join("%g %g %g %g\n" % (press, temp, a, b) for a, b in zip(u, eps))


def main(argv):
    if len(argv) != 5:
        sys.stderr.write("Give parameters: <press> <temp> <spec> <filter>\n")
        return 1
    from jurassic_amd import gpu
    press, temp = float(argv[1]), float(argv[2])
    nuf, filt = read_filter(argv[4])
    tau, nu0, nu1 = read_spectrum(argv[3])
    u, eps = gpu.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
    sys.stdout.write(rows_text(press, temp, u, eps))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
