"""CPU oracle of the SAMPO agent step -- test infrastructure only."""
