"""Writes corner-faulted.data: a compositional deck on a faulted
corner-point grid (COORD / ZCORN from opencaeporo_amd.workloads.
corner_point_arrays: 12 x 10 x 5 cells, leaning pillars, two faults along i
and one along j with throws of up to 1.7 layers), the SPE5 fluid and tables
of ../spe5 (its EGOIL.in, copied beside this deck: INCLUDE paths cannot hold a slash), a water injector and a producer.  A test
input for the deck reader's corner-point path; run from the repository root:
python tests/golden/decks/corner/gen_corner_deck.py"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(HERE)))))

from opencaeporo_amd import workloads  # noqa: E402

NX, NY, NZ = 12, 10, 5


def rows(vals, per=6):
    return "\n".join(" ".join(f"{v:.6f}" for v in vals[i:i + per]) for i in range(0, len(vals), per))


def main():
    coord, zcorn = workloads.corner_point_arrays(NX, NY, NZ, seed=7, nfault=(2, 1), tilt=0.05, jitter=0.1)
    text = f"""-- Compositional deck on a faulted corner-point grid (generated by
-- gen_corner_deck.py; the fluid and tables are those of ../spe5).
RUNSPEC

DIMENS
{NX} {NY} {NZ} /

OIL
WATER
GAS
FIELD

COMPS
6
/

GRID

COORD
{rows(coord)}
/

ZCORN
{rows(zcorn, 8)}
/

EQUALS
'PORO'   0.3   4*  1  {NZ}  /
'PERMX'  300   4*  1  2  /
'PERMX'   60   4*  3  3  /
'PERMX'  150   4*  4  {NZ}  /
'PERMZ'   30   4*  1  {NZ}  /
/

COPY
'PERMX' 'PERMY' 6*  /
/

PROPS

INCLUDE
EGOIL.in
/

ZMFVD
1000.0  0.5 0.03 0.07 0.2 0.15 0.05
10000.0 0.5 0.03 0.07 0.2 0.15 0.05
/

RTEMP
160
/

SWOF
  0.2        0       1       0
  0.2899     0.0022  0.6769  0
  0.3778     0.018   0.4153  0
  0.4667     0.0607  0.2178  0
  0.5556     0.1438  0.0835  0
  0.6444     0.2809  0.0123  0
  0.7        0.4089  0       0
  0.7333     0.4855  0       0
  0.8222     0.7709  0       0
  0.9111     1       0       0
  1          1       0       0
/

SGOF
  0.0      0.0    1.0     0.0
  0.05     0.0    0.88    0.0
  0.0889   0.001  0.7023  0.0
  0.1778   0.01   0.4705  0.0
  0.2667   0.03   0.2963  0.0
  0.3556   0.05   0.1715  0.0
  0.4444   0.1    0.0878  0.0
  0.5333   0.2    0.037   0.0
  0.6222   0.35   0.011   0.0
  0.65     0.39   0.0     0.0
  0.7111   0.56   0.0     0.0
  0.8      1.0    0.0     0.0
/

PVTW
14.7   1.00    3.3E-06      0.7     0.00E-01
/

ROCK
LINEAR01  3990.30 5E-06
/

SOLUTION

EQUIL
8400 4000 9000 0 7000 0 1 1 0  /

SCHEDULE

WELSPECS
'PROD1'   'G'   11   9    1*    'OIL'   /
'INJ1'    'G'    2   2    1*    'WATER' /
/

COMPDAT
'PROD1'   2*   3   5        1*   0.5   3*   /
'INJ1'    2*   1   3        1*   0.5   3*   /
/

WCONPROD
'PROD1'   'OPEN'  'ORAT'  2000    1000    /
/

WCONINJE
'INJ1'    WATER   OPEN    RATE      3000         6000 /
/

TSTEP
10
/
"""
    path = os.path.join(HERE, "corner-faulted.data")
    with open(path, "w") as f:
        f.write(text)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
