"""CPU oracle of the reference collide-stream path (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the product (lb_amd) never imports it.
"""
