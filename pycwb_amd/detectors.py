"""Interferometer site geometry (radians, metres).

The same physical constants as pycwb/constants/detectors.py:1-162 (LAL's
detector table): vertex geodetic latitude/longitude/elevation and, per arm, the
azimuth, altitude and arm midpoint.  Only the sites the likelihood configs use
are listed.
"""

DETECTORS = {
    "H1": {"lat": 0.81079526383, "lon": -2.08405676917, "elevation": 142.554,
           "x": {"alt": -0.00061950000, "az": 5.65487724844, "midpoint": 1997.542},
           "y": {"alt": 0.00001250000, "az": 4.08408092164, "midpoint": 1997.522}},
    "L1": {"lat": 0.53342313506, "lon": -1.58430937078, "elevation": -6.574,
           "x": {"alt": -0.00031210000, "az": 4.40317772346, "midpoint": 1997.575},
           "y": {"alt": -0.00061070000, "az": 2.83238139666, "midpoint": 1997.575}},
    "V1": {"lat": 0.76151183984, "lon": 0.18333805213, "elevation": 51.884,
           "x": {"alt": 0.0, "az": 0.33916285222, "midpoint": 1500.0},
           "y": {"alt": 0.0, "az": 5.05155183261, "midpoint": 1500.0}},
    "I1": {"lat": 0.24841853020, "lon": 1.33401332494, "elevation": 0.0,
           "x": {"alt": 0.0, "az": 1.57079637051, "midpoint": 2000.0},
           "y": {"alt": 0.0, "az": 0.0, "midpoint": 2000.0}},
    "K1": {"lat": 0.6355068497, "lon": 2.396441015, "elevation": 414.181,
           "x": {"alt": 0.0031414, "az": 1.054113, "midpoint": 1513.2535},
           "y": {"alt": -0.0036270, "az": -0.5166798, "midpoint": 1511.611}},
    "G1": {"lat": 0.91184982752, "lon": 0.17116780435, "elevation": 114.425,
           "x": {"alt": 0.0, "az": 1.19360100484, "midpoint": 300.0},
           "y": {"alt": 0.0, "az": 5.83039279401, "midpoint": 300.0}},
}
