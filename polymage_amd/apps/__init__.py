"""The benchmark pipelines written in the PolyMage DSL, with seeded input
generators.  Each module mirrors one of the reference's application files
(sandbox/apps/python/img_proc/*/polymage_*.py) or, where the reference has
no DSL source, the checked-in generated C++ it describes."""
