"""hosono-2dmhd_amd: the hosono-2DMHD per-step update on MI355X.

The directory name carries a hyphen, so the package is imported under the
name ``hosono_2dmhd_amd`` through ``load()`` in bench.py, tests/conftest.py
and __graft_entry__.py (importlib with an explicit location).
"""
