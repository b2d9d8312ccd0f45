"""MI355X-native propagation hot path for Luna-style UPPE/GNLSE simulations."""
