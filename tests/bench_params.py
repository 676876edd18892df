"""The serving stream's seed, read from bench.py without importing it (the
tests draw requests exactly as the benchmark does)."""
import ast
import os

_src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")).read()
SERVE_SEED = next(ast.literal_eval(n.value) for n in ast.parse(_src).body
                  if isinstance(n, ast.Assign) and getattr(n.targets[0], "id", None) == "SERVE_SEED")
