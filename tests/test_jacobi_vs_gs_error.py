"""The error of the GPU's Jacobi update order against the reference's
Gauss-Seidel order (demo4.cpp:223-255), measured with the oracle on the
4,096-particle dam break (tests/jacobi_gs_error.py prints the whole table to
step 1,000).

The GPU step equals the oracle's FP32 Jacobi run bit for bit, so its error
against FP64 is the fp32_jacobi column.  Asserted:
  * FP32 rounding, not the update order, dominates the difference from the
    FP64 run through step 100: the order's own effect (FP64 Jacobi against
    FP64 Gauss-Seidel) is below 1e-9 h through step 20 and below a tenth of
    the rounding error through step 100;
  * the GPU's FP32 error is the reference order's FP32 error (within a factor
    of 1.5 at every mark);
  * the 1e-4 h horizon of the FP32 error lies between steps 13 and 20, and the
    curve stays within a factor of two of the measured values (5.8e-5 h at
    step 10, 2.2e-4 at 20, 1.4e-3 at 50, 5.4e-3 at 100).
"""
from tests.jacobi_gs_error import curves

MEASURED = {10: 5.8e-5, 20: 2.2e-4, 50: 1.4e-3, 100: 5.4e-3}


def test_jacobi_order_error_is_below_fp32_rounding():
    rows = curves((10, 13, 20, 50, 100))
    for m, r in rows.items():
        print(m, r)
    for m in (10, 13, 20):
        assert rows[m]["order"] < 1e-9, rows[m]
    for m, r in rows.items():
        assert r["order"] < 0.1 * r["fp32_jacobi"], (m, r)
        assert r["fp32_jacobi"] <= 1.5 * r["fp32_gs"] and r["fp32_gs"] <= 1.5 * r["fp32_jacobi"], (m, r)
    assert rows[13]["fp32_jacobi"] <= 1e-4 < rows[20]["fp32_jacobi"]
    for m, e in MEASURED.items():
        assert 0.5 * e <= rows[m]["fp32_jacobi"] <= 2.0 * e, (m, rows[m])
