"""CPU oracle package -- test and measurement infrastructure only (never imported by the product)."""
