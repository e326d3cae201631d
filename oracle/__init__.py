"""The oracle: a CPU restatement of Bulbit's path tracer (test infrastructure).

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
package, and only as the checker or the CPU baseline; the product never calls
it.
"""
from .binding import OracleScene, build_oracle, correctly_rounded_libm, load_oracle, load_oracle_from  # noqa: F401
