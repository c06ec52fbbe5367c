"""MI355X-native P3T hot path of GPLUM (soft force, neighbour clusters, batched Hermite)."""
