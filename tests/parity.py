"""The largest relative error per (case, output) that the GPU parity tests of a session measured against the
oracle, with its bar. tests/conftest.py prints them as one table at the end of the run, so the figures appear in
the run's own output; note() returns the error so a test can assert on it in the same expression."""

RESULTS = {}  # (case, output) -> [largest error, bar, checks]


def note(case: str, output: str, err: float, bar: float) -> float:
    r = RESULTS.get((case, output))
    if r is None:
        RESULTS[(case, output)] = [float(err), float(bar), 1]
    else:
        r[0] = max(r[0], float(err))
        r[2] += 1
    return float(err)
