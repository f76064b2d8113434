"""Oracle package: TEST INFRASTRUCTURE ONLY.

Holds the CPU restatement of the reference's flash and FIM assembly (the
checker for the HIP path and the benchmark's CPU baseline) and the recipe that
compiles the reference's own flash into _ref/.  Only tests/, the smoke check
and the benchmark's cpu_baseline leg may import anything from here.
"""
