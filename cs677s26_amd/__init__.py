"""MI355X-native Viola-Jones hot path (integrals, Haar stump search, cascade detect)."""
