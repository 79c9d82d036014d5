"""CPU oracle for the JST + Spalart-Allmaras loop (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import this
package, and only as a checker or as the timed CPU baseline -- never as the product path.
"""
